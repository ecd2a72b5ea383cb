/*
 * shape_atrophy.h -- C ABI of the MI355X (gfx950) forward shape-evolution
 * engine.
 *
 * The reference exposes this path as four MATLAB MEX functions, called only
 * by CUDA/deformation.m and CUDA/computeObjfcn.m:
 *
 *   deformation_2D_frame / deformation_3D_frame       (2D|3D_frame/deformation.cu:14)
 *   computeObjfcn_2D_frame / computeObjfcn_3D_frame   (2D|3D_frame/computeObjfcn.cu:14)
 *
 * Each takes the mesh index tables (elmObj), the initial landmarks, the model
 * parameters and a matrix of candidate tau variables, and loops over the
 * candidate columns (deformation.cu:152-156, computeObjfcn.cu:162-166).  The
 * structs below carry exactly those positional arguments (the argument
 * number of each field is given) and the two entry points
 * sa_deformation / sa_compute_objfcn replace the two MEX families one to one;
 * the dimension is a field instead of part of the function name.
 *
 * Layouts are the reference's: every table is column-major (MATLAB) with
 * 0-based int32 indices, landmark matrices are lmkNum x dim column-major,
 * lmkStk is (lmkNum*dim) x timeNum, tauMat lmkNum x timeNum.
 *
 * Pointers are host pointers unless `on_device` is set in sa_mesh, in which
 * case every array of sa_mesh, sa_target and lmkIniMat is a device pointer
 * (what mxGPUGetDataReadOnly yields, deformation.cu:67-77).  Outputs are
 * always host memory (the MEX copies results back, deformation.cu:163-164).
 *
 * Return value: 0 on success, a negative SA_E* code on a hard error (the
 * mexErrMsgIdAndTxt cases of deformation.cu:116-129); sa_last_error() then
 * describes it.  A candidate whose velocity solve fails gets objective
 * DBL_MAX and its trajectory stops, as deform.cu:68-73 does.
 */
#ifndef SHAPE_ATROPHY_H
#define SHAPE_ATROPHY_H

#ifdef __cplusplus
extern "C" {
#endif

#define SA_OK 0
#define SA_E_ARG -1        /* bad sizes / null pointers / unsupported option */
#define SA_E_ALLOC -2      /* device allocation failed ("objfcn*:cudaMalloc") */
#define SA_E_DEVICE -3     /* HIP runtime error */
#define SA_E_NOTFOUND -4   /* unknown session / request handle */

/* elasticity model: 'frame' = the reference's anisotropic frame operator;
   'isotropic' = per-layer Lame pairs (lambda, mu) scaled by (1 - 0.5 u), the
   branch CUDA/deformation.m:20-24 reserves but leaves empty. */
#define SA_ELA_FRAME 0
#define SA_ELA_ISOTROPIC 1

/* arithmetic of the velocity solve */
#define SA_PREC_FP64 0     /* plain FP64 everywhere */
#define SA_PREC_MIXED 1    /* FP32 kernel products inside FP64 iterative refinement */

typedef struct sa_mesh {
    int dim;               /* 2 or 3 */
    int on_device;         /* 0: host pointers, 1: device pointers */
    int lmkNum;            /* rows of lmkIniMat (arg 2) */
    int elmNum;            /* rows of elmVtxMat (arg 3) */
    int vtxElmCols;        /* columns of vtxElmMat (arg 4) = 1 + 2 * max adjacency */
    int bdrElmNum;         /* rows of bdrVtxMat (arg 7) */
    int vtxBdrCols;        /* columns of vtxBdrMat (arg 8) */
    int btmElmNum;         /* rows of btmVtxMat (arg 9) */
    int btmLmkNum;         /* rows of vtxBtmMat (arg 10) */
    int vtxBtmCols;        /* columns of vtxBtmMat */
    int nzrNum;            /* rows of femVtxMat (arg 11) */
    int femIfoCols;        /* columns of femIfoMat (arg 12) */
    int vfdElmNum;         /* rows of vfdVtxMat (computeObjfcn arg 13); 0 for deformation */
    const int *elmVtxMat;  /* arg 3  elmNum x (dim+1) */
    const int *vtxElmMat;  /* arg 4  lmkNum x vtxElmCols */
    const int *frmVtxMat;  /* arg 5  tanVtxMat (2D, elmNum x 4) | nmlVtxMat (3D, elmNum x 6) */
    const int *tsvVtxMat;  /* arg 6  elmNum x 4 (2D) | 6 (3D) */
    const int *bdrVtxMat;  /* arg 7  (read only with sa_params.bdrYank; accepted for call compatibility) */
    const int *vtxBdrMat;  /* arg 8  (read only with sa_params.bdrYank) */
    const int *btmVtxMat;  /* arg 9  btmElmNum x dim */
    const int *vtxBtmMat;  /* arg 10 btmLmkNum x vtxBtmCols */
    const int *femVtxMat;  /* arg 11 nzrNum x 2, sorted (row, col) */
    const int *femIfoMat;  /* arg 12 nzrNum x femIfoCols */
    const int *vfdVtxMat;  /* computeObjfcn arg 13, vfdElmNum x dim */
    const int *elmLyrVec;  /* isotropic model only: material layer of each element (may be NULL) */
    int ctlNum;            /* > 0: RKHS kernel on these control landmarks (build-defined form) */
    const int *ctlIdxVec;  /* ctlNum landmark indices */
} sa_mesh;

typedef struct sa_params {
    int elaType;                 /* SA_ELA_FRAME | SA_ELA_ISOTROPIC */
    double ldmWgt;               /* arg 13 (deformation numbering) */
    int knlOrder;                /* arg 14: -1 Gaussian, 0..4 Matern */
    double knlWidth;             /* arg 15 */
    double knlEps;               /* arg 16 */
    const double *modVec;        /* arg 17: frame 3 (2D) / 4 (3D) moduli; isotropic 2 per layer */
    int modNum;                  /* length of modVec */
    double spdTanVal;            /* arg 18 */
    double spdTsvVal;            /* arg 19 */
    const double *ynkActPrmVec;  /* arg 20: tauMin, tauMax, actMax, dWidth */
    const double *reaActPrmVec;  /* arg 21: tauMin, tauMax, actMax, dWidth, hat */
    double btmWgt;               /* arg 22 */
    double timeStp;              /* arg 23 */
    int timeNum;                 /* arg 24 = floor(timeEnd / timeStp) + 1 */
    int cgItrMax;                /* arg 25 */
    double cgTolSqu;             /* arg 26 = cgTolVal^2 */
    int precision;               /* SA_PREC_FP64 | SA_PREC_MIXED */
    double rdTolVal;             /* relative residual of the sparse reaction-diffusion solve (0 -> 1e-13) */
    int bdrYank;                 /* 1: add the boundary yank the reference keeps commented out
                                    (computeExternalYank.cu:49-84 | 3D :58-99, gathered through
                                    vtxBdrMat :106-116 | :125-135); needs bdrVtxMat/vtxBdrMat.
                                    0: as the reference runs. */
    int velPrec;                 /* preconditioner of the velocity CG: SA_VELPREC_DEFAULT (the
                                    reference's K preconditioning; control points: (K_cc + eps I)^-1
                                    with a spectral term) | SA_VELPREC_JACOBI (control-point form
                                    only: the diagonal of the control-point system) */
} sa_params;

#define SA_VELPREC_DEFAULT 0
#define SA_VELPREC_JACOBI 1

typedef struct sa_target {
    int tgtElmNum;               /* rows of the target matrices */
    const double *cenPosMat;     /* computeObjfcn arg 14, tgtElmNum x dim */
    const double *uniDirMat;     /* arg 15, tgtElmNum x dim */
    const double *elmVolVec;     /* arg 16, tgtElmNum */
    char cenKnlType;             /* arg 17: 'G' Gaussian | 'C' Cauchy */
    double cenKnlWidth;          /* arg 18 */
    char dirKnlType;             /* arg 19: 'B' Binet | 'L' linear | 'O' | 'U' */
    double dirKnlWidth;          /* arg 20 */
} sa_target;

/* per-solve counters */
typedef struct sa_stats {
    int status;                  /* 0 ok, 1 velocity solve failed (objective = DBL_MAX) */
    int steps;                   /* time steps completed */
    int pcgIters;                /* total velocity-solve iterations */
    int rdIters;                 /* total reaction-diffusion iterations */
    int refineSweeps;            /* outer FP64 refinement sweeps (mixed precision) */
    double seconds;              /* wall time of the solve */
    double kmatvecMs;            /* with profiling on: summed hipEvent time of the kernel-matrix products */
    int kmatvecCount;            /*   ... and their number (CG iterations that ran one) */
    double operatorMs;           /* summed time of the elastic operator applies (element + gather kernels) */
    int operatorCount;
    int fp64Steps;               /* mixed precision: steps whose refinement stalled and were finished in FP64 */
    int failReason;              /* status 1: SA_FAIL_* */
    double varifoldMs;           /* with profiling on: hipEvent time of the objective's varifold row-sum launch */
    long long varifoldPairs;     /*   ... the facet pairs it summed (F^2 + F Ft, + Ft^2 on a session's first) */
    int varifoldMfma;            /*   ... 1 when the pair contractions ran on the matrix cores */
    long long deviceBytes;       /* device memory the session holds (= sa_session_footprint of its mesh) */
    int factorizations;          /* control-point K_cc factorisations (preconditioner rebuilds) of the run */
} sa_stats;

/* sa_stats.failReason */
#define SA_FAIL_BREAKDOWN 1      /* velocity CG breakdown (zero curvature or zero r.z) */
#define SA_FAIL_BUDGET    2      /* velocity CG iteration budget (cgItrMax) spent */
#define SA_FAIL_CHOLESKY  3      /* control-point K_cc not positive definite (Cholesky) */
#define SA_FAIL_TRTRI     4      /* inverse of the Cholesky factor failed */
#define SA_FAIL_RD        5      /* reaction-diffusion CG did not converge */
#define SA_FAIL_REFINE    6      /* not produced: a mixed-precision step whose refinement stalls
                                    is finished by the FP64 form, which reports its own reason */

/* deformation_{2D,3D}_frame (deformation.cu:14-199).
   tauVarMat: tauVarNum x candNum (arg 0); tauPrmVec (arg 1); lmkIniMat (arg 2).
   Outputs: lmkStk (lmkNum*dim x timeNum) and tauMat (lmkNum x timeNum) of the
   last candidate, objVec (candNum; 0, or DBL_MAX where the solve failed).
   Any output pointer may be NULL. */
int sa_deformation(const sa_mesh *mesh, const sa_params *prm, const double *tauVarMat, int tauVarNum,
                   int candNum, const double *tauPrmVec, int tauPrmNum, const double *lmkIniMat,
                   double *lmkStk, double *tauMat, double *objVec);

/* computeObjfcn_{2D,3D}_frame (computeObjfcn.cu:14-212): the varifold
   distance of the evolved surface to the target for each candidate.
   dqObjMat (nullable, lmkNum x dim) receives the gradient of the last
   candidate's objective with respect to its final landmark positions (the
   compiled-but-unused gradient overload of varifold.cu:545-576|577-608). */
int sa_compute_objfcn(const sa_mesh *mesh, const sa_target *tgt, const sa_params *prm, const double *tauVarMat,
                      int tauVarNum, int candNum, const double *tauPrmVec, int tauPrmNum,
                      const double *lmkIniMat, double *objVec, double *dqObjMat);

/* Device-resident session: the mesh, target and per-step state stay in HBM
   across candidates (the MEX re-allocates its arena on every call).  No
   counterpart in the reference; it is the form the batch and serving layers
   use. */
typedef struct sa_session sa_session;
int sa_session_create(const sa_mesh *mesh, const sa_target *tgt /* nullable */, const sa_params *prm,
                      const double *lmkIniMat, sa_session **out);
/* one candidate; any output may be NULL; returns 0 or an SA_E* code */
int sa_session_run(sa_session *s, const double *tauVarVec, int tauVarNum, const double *tauPrmVec,
                   int tauPrmNum, double *obj, double *lmkStk, double *tauMat, double *dqObjMat,
                   sa_stats *stats);
void sa_session_destroy(sa_session *s);
/* time the kernel-matrix product and the elastic operator of every CG
   iteration with hipEvents on the session's stream (reported in sa_stats) */
int sa_session_set_profiling(sa_session *s, int on);
/* run only the first timeNum time levels (<= the session's timeNum) from now on */
int sa_session_set_time_num(sa_session *s, int timeNum);
/* new scalar parameters (moduli, diffusion speeds, activities, weights,
   kernel width, time step, CG controls) for the following runs; sizes, the
   elasticity model, kernel order and precision stay those of the session */
int sa_session_set_params(sa_session *s, const sa_params *prm);
/* Diagnostic probes (no counterpart in the reference): evaluate one operator
   of the path at the initial state of a candidate (q0 and tau0 from
   tauVarVec) so each kernel family can be checked against an oracle alone.
     SA_PROBE_ELASTIC   out = E(q0) in                       (lmkNum x dim)
     SA_PROBE_KMATVEC   out = (K(q0) + knlEps I) in          (lmkNum x dim)
     SA_PROBE_YANK      out = [f(q0, tau0) (lmkNum x dim), tau0 (lmkNum)]
     SA_PROBE_FEM       out = [M(q0) (nzrNum), M + dt K_D (nzrNum)] in femVtxMat order
     SA_PROBE_GEOMETRY  out = per-element [Q^-1 (dim^2), det Q, frame 1 (dim), frame 2 (dim)] x elmNum
     SA_PROBE_VARIFOLD  in = landmarks (lmkNum x dim); out = [objective, gradient (lmkNum x dim)]
                        (the objective by the kernel a run uses, the gradient by the FP64 VALU kernel)
     SA_PROBE_PENALTY   out = P(q0) in, the bottom penalty without btmWgt (lmkNum x dim)
     SA_PROBE_KXC       out = K_xc(q0) in: in ctlNum x dim, out lmkNum x dim  (control-point
     SA_PROBE_KCX       out = K_cx(q0) in: in lmkNum x dim, out ctlNum x dim   kernel products
                        as the session's inner solve forms them, without knlEps)
     SA_PROBE_JACOBI    out = the diagonal of the control-point system at q0, tau0 (ctlNum x dim;
                        needs a session with velPrec = SA_VELPREC_JACOBI)
     SA_PROBE_KCC_CHOL  out = L, the Cholesky factor of K_cc(q0) + knlEps I as the preconditioner
                        build forms it (ctlNum x ctlNum column-major, upper triangle zero) */
#define SA_PROBE_ELASTIC 1
#define SA_PROBE_KMATVEC 2
#define SA_PROBE_YANK 3
#define SA_PROBE_FEM 4
#define SA_PROBE_GEOMETRY 5
#define SA_PROBE_VARIFOLD 6
#define SA_PROBE_PENALTY 7
#define SA_PROBE_KXC 8
#define SA_PROBE_KCX 9
#define SA_PROBE_JACOBI 10
#define SA_PROBE_KCC_CHOL 12
int sa_session_probe(sa_session *s, int what, const double *tauVarVec, int tauVarNum, const double *tauPrmVec,
                     int tauPrmNum, const double *in, double *out);
/* bytes of device memory a session of this mesh holds */
long long sa_session_footprint(const sa_mesh *mesh, const sa_target *tgt, const sa_params *prm);

/* Tile table (no counterpart in the reference; for the offline sweep that
   writes it).  The stored kernel products launch about `target` blocks (row
   panels x column slices); the targets come from tiles.txt beside the
   library, one line per product, element type and measured shape:
     <xc|cx|cc|kk> <f32|f64> <rows> <cols> <target>
   (xc: K_xc, landmarks x control points; cx: K_cx; cc: the control-point
   preconditioner product; kk: K over all landmarks).  A session takes the
   entry of its product and type nearest its shape.  SA_TILE_TABLE names
   another file; SA_KS_TARGET ("N" or "xc=N,cx=M,...") overrides the table.
   A line "op f64 <elmNum> <lmkNum> <shape>" sets the launch shape (0-3) of
   the fused operator apply for meshes near that size (SA_OPF_CFG overrides).
   sa_tile_target returns the target a new session would take (0: no entry,
   the built-in default applies; -1: bad argument); for product "op" (rows =
   elmNum, cols = lmkNum) the launch shape a session would take.
   sa_session_time_product times one product as a solve launches it (product
   kernel and slice sum) at q0: the mean over reps launches after a warm-up,
   in microseconds; what = SA_PROBE_KXC, SA_PROBE_KCX, SA_PROBE_KMATVEC (the
   all-landmark form), SA_TIME_PRECOND, or SA_PROBE_ELASTIC (the operator
   apply (E + btmWgt P) x at the geometry of the last run's or probe's start). */
/* The vertex chunks of the fused operator apply (one launch: element forces
   in LDS, summed per vertex in vtxElmMat order), built on the host from a
   host mesh and positions as a session builds them, and checked against the
   invariants the kernel's indexing relies on (no counterpart in the
   reference; for tests).  chunks: launch blocks; elementEntries: element
   evaluations per apply (at least elmNum); haloEntries: vertices staged in
   LDS per apply.  SA_E_ARG when a check fails or a vertex's elements pass
   a chunk's capacity (such meshes keep the two-launch apply). */
int sa_op_chunk_check(const sa_mesh *mesh, const double *lmkIniMat, int *chunks, long long *elementEntries,
                      long long *haloEntries);
#define SA_TIME_PRECOND 11
int sa_tile_target(const char *product, int f32, long long rows, long long cols);
int sa_session_time_product(sa_session *s, int what, int reps, double *usPerProduct);

/* Request server (no counterpart in the reference): solves of registered
   meshes run on worker threads, each request on a device-resident session of
   its mesh with its own HIP stream, so requests overlap on the device.
   Requests are admitted in arrival order as workers and device memory free
   up (memBudgetBytes <= 0: 80% of the free device memory); idle sessions of
   other meshes are released, least recently used first, to make room.  A
   request may carry its own scalar parameters (NULL: the mesh's). */
typedef struct sa_server sa_server;
int sa_server_create(int workers, long long memBudgetBytes, sa_server **out);
void sa_server_destroy(sa_server *sv);
int sa_server_add_mesh(sa_server *sv, const sa_mesh *mesh, const sa_target *tgt /* nullable */,
                       const sa_params *prm, const double *lmkIniMat, int *meshId);
int sa_server_submit(sa_server *sv, int meshId, const double *tauVarVec, int tauVarNum, const double *tauPrmVec,
                     int tauPrmNum, const sa_params *prm /* nullable */, long long *reqId);
/* blocks until the request finished; returns its status (SA_OK or SA_E_*);
   times = [submitted, started, finished] in seconds since the server started */
int sa_server_wait(sa_server *sv, long long reqId, double *obj, sa_stats *stats, double *times);
int sa_server_info(sa_server *sv, long long *memUsed, int *sessions, int *queued, int *running);

/* The varifold objective of computeObjfcn (varifold.cu value kernels, dsum.cu)
   for landmarks, facets and target already in device memory, allocating
   nothing.  No counterpart in the reference: it is for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving or
   fitting loops).  It computes what sa_session_probe(SA_PROBE_VARIFOLD)
   returns as its objective, with the same kernels, but keeps its
   intermediate values (facet centres, directions and measures, pair-sum
   rows) in a fixed device array of the library (8 MiB per device, held
   until the process ends) instead of a session.
     lmkMat     device, lmkNum x dim (column-major, as lmkIniMat)
     vfdVtxMat  device, vfdElmNum x dim int32 (computeObjfcn arg 13)
     tgt        host struct whose cenPosMat / uniDirMat / elmVolVec are device pointers
     objDev     device, 1 double: the objective (NaN when a facet index is outside [0, lmkNum))
     stream     hipStream_t of the launches (NULL: the null stream); the call only enqueues work
   Limit: (2 dim + 2) vfdElmNum + tgtElmNum + 3 <= 1,048,576, else SA_E_ARG.
   Every call on a device shares that device's array: calls that overlap on
   one device (other streams or threads) race and must be serialised by the
   caller. */
int sa_varifold_objective_noalloc(int dim, int lmkNum, const double *lmkMat, int vfdElmNum, const int *vfdVtxMat,
                                  const sa_target *tgt, double *objDev, void *stream);

const char *sa_last_error(void);
int sa_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* SHAPE_ATROPHY_H */
