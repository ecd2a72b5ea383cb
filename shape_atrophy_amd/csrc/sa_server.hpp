// Request server over device-resident sessions (sa_server_* in the C ABI).
//
// A request is one forward solve + objective of a registered mesh with its
// own tau variables and, optionally, its own scalar parameters (the
// parameter candidates a model fit evaluates).  Worker threads take requests
// in arrival order; each runs on a session of the request's mesh, and every
// session has its own HIP stream, so requests placed on different sessions
// overlap on the device.  Sessions are created on demand and kept (mesh and
// solver state stay in HBM) while the device-memory budget allows; to admit a
// request whose mesh has no idle session, idle sessions of other meshes are
// destroyed, least recently used first.  The head request waits until a
// worker and enough memory are free (admission by HBM footprint).
#pragma once
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace sa {

struct MeshCopy {
    sa_mesh m{};
    sa_target t{};
    bool hasT = false;
    sa_params p{};
    std::vector<int> elmVtx, vtxElm, frm, tsv, bdr, vtxBdr, btm, vtxBtm, femVtx, femIfo, vfd, lyr, ctl;
    std::vector<double> lmk, tcen, tdir, tvol, mod, ynk, rea;
    long long footprint = 0;

    template <typename T>
    static std::vector<T> cp(const T *src, size_t n) {
        if (!src || n == 0) return {};
        return std::vector<T>(src, src + n);
    }

    MeshCopy(const sa_mesh &mm, const sa_target *tt, const sa_params &pp, const double *lmkIni) {
        if (mm.on_device) throw SaError(SA_E_ARG, "the server takes host arrays");
        const int D = mm.dim;
        m = mm;
        p = pp;
        elmVtx = cp(mm.elmVtxMat, (size_t)mm.elmNum * (D + 1));
        vtxElm = cp(mm.vtxElmMat, (size_t)mm.lmkNum * mm.vtxElmCols);
        frm = cp(mm.frmVtxMat, (size_t)mm.elmNum * (D == 2 ? 4 : 6));
        tsv = cp(mm.tsvVtxMat, (size_t)mm.elmNum * (D == 2 ? 4 : 6));
        bdr = cp(mm.bdrVtxMat, (size_t)mm.bdrElmNum * D);
        vtxBdr = cp(mm.vtxBdrMat, (size_t)mm.lmkNum * mm.vtxBdrCols);
        btm = cp(mm.btmVtxMat, (size_t)mm.btmElmNum * D);
        vtxBtm = cp(mm.vtxBtmMat, (size_t)mm.btmLmkNum * mm.vtxBtmCols);
        femVtx = cp(mm.femVtxMat, (size_t)mm.nzrNum * 2);
        femIfo = cp(mm.femIfoMat, (size_t)mm.nzrNum * mm.femIfoCols);
        vfd = cp(mm.vfdVtxMat, (size_t)mm.vfdElmNum * D);
        lyr = cp(mm.elmLyrVec, (size_t)mm.elmNum);
        ctl = cp(mm.ctlIdxVec, (size_t)std::max(0, mm.ctlNum));
        lmk = cp(lmkIni, (size_t)mm.lmkNum * D);
        mod = cp(pp.modVec, (size_t)pp.modNum);
        ynk = cp(pp.ynkActPrmVec, 4);
        rea = cp(pp.reaActPrmVec, 5);
        if (tt) {
            hasT = true;
            t = *tt;
            tcen = cp(tt->cenPosMat, (size_t)tt->tgtElmNum * D);
            tdir = cp(tt->uniDirMat, (size_t)tt->tgtElmNum * D);
            tvol = cp(tt->elmVolVec, (size_t)tt->tgtElmNum);
        }
        m.elmVtxMat = elmVtx.data(); m.vtxElmMat = vtxElm.data(); m.frmVtxMat = frm.data();
        m.tsvVtxMat = tsv.data(); m.bdrVtxMat = bdr.empty() ? nullptr : bdr.data();
        m.vtxBdrMat = vtxBdr.empty() ? nullptr : vtxBdr.data();
        m.btmVtxMat = btm.empty() ? nullptr : btm.data(); m.vtxBtmMat = vtxBtm.empty() ? nullptr : vtxBtm.data();
        m.femVtxMat = femVtx.data(); m.femIfoMat = femIfo.data(); m.vfdVtxMat = vfd.empty() ? nullptr : vfd.data();
        m.elmLyrVec = lyr.empty() ? nullptr : lyr.data(); m.ctlIdxVec = ctl.empty() ? nullptr : ctl.data();
        p.modVec = mod.data(); p.ynkActPrmVec = ynk.data(); p.reaActPrmVec = rea.data();
        if (hasT) { t.cenPosMat = tcen.data(); t.uniDirMat = tdir.data(); t.elmVolVec = tvol.data(); }
        footprint = sa_session_footprint(&m, hasT ? &t : nullptr, &p);
        // the sizing pass validates the tables: a mesh no session could be
        // built for is refused at registration
        if (footprint < 0) throw SaError(SA_E_ARG, sa_last_error());
    }
};

struct Request {
    long long id = 0;
    int mesh = 0;
    std::vector<double> tv, tp;
    bool hasPrm = false;
    sa_params prm{};
    std::vector<double> mod, ynk, rea;
    int state = 0;          // 0 queued, 1 running, 2 done
    int rc = 0;
    std::string err;
    double obj = 0.0;
    sa_stats st{};
    double tSubmit = 0, tStart = 0, tEnd = 0;
    int session = -1;
};

struct Slot {
    sa_session *s = nullptr;
    int mesh = -1;
    bool busy = false;
    long long bytes = 0;
    double lastUse = 0;
};

class Server {
  public:
    Server(int workers, long long budget) : budget_(budget) {
        t0_ = std::chrono::steady_clock::now();
        for (int i = 0; i < std::max(1, workers); ++i) threads_.emplace_back([this] { work(); });
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
        for (auto &sl : slots_) if (sl.s) sa_session_destroy(sl.s);
    }

    int add_mesh(const sa_mesh &m, const sa_target *t, const sa_params &p, const double *lmk) {
        auto mc = std::make_unique<MeshCopy>(m, t, p, lmk);
        std::lock_guard<std::mutex> g(mu_);
        meshes_.push_back(std::move(mc));
        return (int)meshes_.size() - 1;
    }

    long long submit(int mesh, const double *tv, int nv, const double *tp, int np, const sa_params *prm) {
        std::lock_guard<std::mutex> g(mu_);
        if (mesh < 0 || mesh >= (int)meshes_.size()) throw SaError(SA_E_NOTFOUND, "unknown mesh id");
        if (!tv || !tp || nv <= 0 || np <= 0) throw SaError(SA_E_ARG, "bad tau arguments");
        Request r;
        r.id = ++lastId_;
        r.mesh = mesh;
        r.tv.assign(tv, tv + nv);
        r.tp.assign(tp, tp + np);
        if (prm) {
            r.hasPrm = true;
            r.prm = *prm;
            if (!prm->modVec || !prm->ynkActPrmVec || !prm->reaActPrmVec) throw SaError(SA_E_ARG, "null parameter vector");
            r.mod.assign(prm->modVec, prm->modVec + prm->modNum);
            r.ynk.assign(prm->ynkActPrmVec, prm->ynkActPrmVec + 4);
            r.rea.assign(prm->reaActPrmVec, prm->reaActPrmVec + 5);
        }
        r.tSubmit = now();
        reqs_[r.id] = std::move(r);
        queue_.push_back(lastId_);
        cv_.notify_all();
        return lastId_;
    }

    // blocks until the request has finished
    Request wait(long long id) {
        std::unique_lock<std::mutex> g(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) throw SaError(SA_E_NOTFOUND, "unknown request id");
        cv_.wait(g, [&] { return reqs_[id].state == 2; });
        Request r = reqs_[id];
        reqs_.erase(id);
        return r;
    }

    void info(long long *memUsed, int *sessions, int *queued, int *running) {
        std::lock_guard<std::mutex> g(mu_);
        if (memUsed) *memUsed = memUsed_;
        if (sessions) *sessions = (int)std::count_if(slots_.begin(), slots_.end(), [](const Slot &s) { return s.s != nullptr; });
        if (queued) *queued = (int)queue_.size();
        if (running) *running = running_;
    }

  private:
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }

    // under the lock: choose a session for the head request.  Returns a
    // slot, -1 (not now: wait for a request to finish or an eviction to
    // complete), -2 (can never fit) or -3 (idle sessions must go first: they
    // are moved to `evict`, their memory still counted in memUsed_ until the
    // caller has destroyed them outside the lock and released it)
    int place(const Request &r, bool &create, std::vector<Slot> &evict) {
        create = false;
        for (int i = 0; i < (int)slots_.size(); ++i)
            if (slots_[i].s && !slots_[i].busy && slots_[i].mesh == r.mesh) return i;
        long long need = meshes_[r.mesh]->footprint;
        if (need > budget_) return -2;   // can never fit
        if (memUsed_ + need > budget_) {
            if (evicting_) return -1;    // another worker is releasing memory
            long long freed = 0;
            while (memUsed_ - freed + need > budget_) {
                int lru = -1;
                for (int i = 0; i < (int)slots_.size(); ++i)
                    if (slots_[i].s && !slots_[i].busy && (lru < 0 || slots_[i].lastUse < slots_[lru].lastUse))
                        lru = i;
                if (lru < 0) break;      // not enough idle sessions: wait for running requests
                freed += slots_[lru].bytes;
                evict.push_back(slots_[lru]);
                slots_[lru] = Slot{};
            }
            if (evict.empty()) return -1;
            return -3;
        }
        int idx = -1;
        for (int i = 0; i < (int)slots_.size(); ++i)
            if (!slots_[i].s && !slots_[i].busy) { idx = i; break; }
        if (idx < 0) { slots_.push_back(Slot{}); idx = (int)slots_.size() - 1; }
        slots_[idx].mesh = r.mesh;
        slots_[idx].bytes = need;
        memUsed_ += need;          // reserved before the session exists
        create = true;
        return idx;
    }

    void work() {
        std::unique_lock<std::mutex> g(mu_);
        while (true) {
            int slot = -1;
            bool create = false;
            long long id = 0;
            std::vector<Slot> evict;
            cv_.wait(g, [&] {
                if (stop_) return true;
                if (queue_.empty()) return false;
                id = queue_.front();
                slot = place(reqs_[id], create, evict);
                return slot != -1;
            });
            if (stop_) {
                // (sessions taken for eviction are still destroyed)
                g.unlock();
                for (auto &sl : evict) sa_session_destroy(sl.s);
                return;
            }
            if (slot == -3) {
                // destroy the evicted idle sessions without holding the lock
                // (their hipFree waits for the device), then release their
                // memory and place the head request again
                evicting_ = true;
                g.unlock();
                long long freed = 0;
                for (auto &sl : evict) {
                    sa_session_destroy(sl.s);
                    freed += sl.bytes;
                }
                g.lock();
                memUsed_ -= freed;
                evicting_ = false;
                cv_.notify_all();
                continue;
            }
            queue_.pop_front();
            Request &r = reqs_[id];
            if (slot == -2) {
                r.rc = SA_E_ALLOC;
                r.err = "request does not fit the device-memory budget";
                r.state = 2;
                r.tStart = r.tEnd = now();
                cv_.notify_all();
                continue;
            }
            slots_[slot].busy = true;
            r.state = 1;
            r.session = slot;
            r.tStart = now();
            ++running_;
            MeshCopy *mc = meshes_[r.mesh].get();
            sa_session *sess = slots_[slot].s;
            Request job = r;   // private copy for the unlocked part
            g.unlock();
            int rc = SA_OK;
            std::string err;
            double obj = 0.0;
            sa_stats st{};
            if (create) {
                rc = sa_session_create(&mc->m, mc->hasT ? &mc->t : nullptr, &mc->p, mc->lmk.data(), &sess);
                if (rc != SA_OK) { err = sa_last_error(); sess = nullptr; }
            }
            if (rc == SA_OK) {
                sa_params prm = mc->p;
                if (job.hasPrm) {
                    prm = job.prm;
                    prm.modVec = job.mod.data();
                    prm.ynkActPrmVec = job.ynk.data();
                    prm.reaActPrmVec = job.rea.data();
                }
                rc = sa_session_set_params(sess, &prm);
                if (rc == SA_OK)
                    rc = sa_session_run(sess, job.tv.data(), (int)job.tv.size(), job.tp.data(), (int)job.tp.size(),
                                        &obj, nullptr, nullptr, nullptr, &st);
                if (rc != SA_OK) err = sa_last_error();
            }
            g.lock();
            --running_;
            Request &rr = reqs_[id];
            rr.rc = rc;
            rr.err = err;
            rr.obj = obj;
            rr.st = st;
            rr.tEnd = now();
            rr.state = 2;
            Slot &sl = slots_[slot];
            sl.busy = false;
            sl.lastUse = rr.tEnd;
            if (create) {
                if (sess) sl.s = sess;
                else { memUsed_ -= sl.bytes; sl = Slot{}; }
            }
            cv_.notify_all();
        }
    }

    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<std::unique_ptr<MeshCopy>> meshes_;
    std::deque<long long> queue_;
    std::unordered_map<long long, Request> reqs_;
    std::vector<Slot> slots_;
    std::vector<std::thread> threads_;
    long long budget_, memUsed_ = 0, lastId_ = 0;
    int running_ = 0;
    bool stop_ = false;
    bool evicting_ = false;   // a worker is destroying evicted sessions
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace sa
