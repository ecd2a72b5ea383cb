// The request server (shape_atrophy_amd/csrc/sa_server.hpp) built against a
// stub engine on the CPU, for ThreadSanitizer / AddressSanitizer runs of
// submit, wait and eviction.  The stub sessions check their own state
// (double create, use after destroy), sleep a little per run, and the
// process checks that the bytes of live sessions never exceed the server's
// budget and that every result is the pure function of its request.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>

#include "shape_atrophy.h"

namespace sa {
struct SaError : std::runtime_error {
    int code;
    SaError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
}  // namespace sa

namespace {
thread_local std::string g_err;
std::atomic<long long> g_live{0}, g_peak{0}, g_budget{0};
std::atomic<int> g_sessions{0}, g_violations{0};
constexpr unsigned kMagic = 0x5e5510a1u;

long long stub_bytes(const sa_mesh *m) { return 1000LL * m->lmkNum; }
double stub_obj(const double *tv, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += (i + 1) * tv[i] * tv[i];
    return s;
}
}  // namespace

struct sa_session {
    unsigned magic = kMagic;
    long long bytes = 0;
    int lmk = 0;
    sa_params prm{};
};

extern "C" {
const char *sa_last_error(void) { return g_err.c_str(); }
long long sa_session_footprint(const sa_mesh *m, const sa_target *, const sa_params *) {
    if (!m || m->lmkNum <= 0) {
        g_err = "invalid mesh";
        return -1;
    }
    return stub_bytes(m);
}
int sa_session_create(const sa_mesh *m, const sa_target *, const sa_params *p, const double *lmk, sa_session **out) {
    if (!m || !p || !lmk || !out) return SA_E_ARG;
    auto *s = new sa_session;
    s->bytes = stub_bytes(m);
    s->lmk = m->lmkNum;
    s->prm = *p;
    const long long live = g_live += s->bytes;
    long long peak = g_peak.load();
    while (live > peak && !g_peak.compare_exchange_weak(peak, live)) {}
    if (live > g_budget.load()) g_violations++;
    g_sessions++;
    *out = s;
    return SA_OK;
}
int sa_session_set_params(sa_session *s, const sa_params *p) {
    if (!s || s->magic != kMagic || !p) { g_violations++; return SA_E_NOTFOUND; }
    s->prm = *p;
    return SA_OK;
}
int sa_session_run(sa_session *s, const double *tv, int nv, const double *, int, double *obj, double *, double *,
                   double *, sa_stats *st) {
    if (!s || s->magic != kMagic) { g_violations++; return SA_E_NOTFOUND; }
    std::this_thread::sleep_for(std::chrono::microseconds(100 + (s->lmk % 7) * 50));
    if (obj) *obj = stub_obj(tv, nv) * s->prm.ldmWgt;
    if (st) { *st = sa_stats{}; st->steps = s->prm.timeNum - 1; st->deviceBytes = s->bytes; }
    return SA_OK;
}
void sa_session_destroy(sa_session *s) {
    if (!s) return;
    if (s->magic != kMagic) g_violations++;
    s->magic = 0;
    g_live -= s->bytes;
    g_sessions--;
    delete s;
}
}

#include "../../shape_atrophy_amd/csrc/sa_server.hpp"

int main(int argc, char **argv) {
    const int nreq = argc > 1 ? std::atoi(argv[1]) : 400;
    const long long budget = 2 * 1000LL * 3000 + 500;   // room for about two of the larger meshes
    g_budget = budget;
    int failures = 0;
    {
        sa::Server srv(4, budget);
        // five meshes (landmark counts 500 .. 3000) and one that never fits
        std::vector<int> sizes = {500, 1200, 3000, 2000, 800, 9000};
        std::vector<int> ids;
        std::vector<int> elm(4, 0), vtx(4, 0);
        std::vector<double> lmk(3 * 9000, 0.0), mod(4, 1.0), ynk(4, 0.0), rea(5, 0.0);
        for (int v : sizes) {
            sa_mesh m{};
            m.dim = 3;
            m.lmkNum = v;
            m.elmVtxMat = elm.data();
            m.vtxElmMat = vtx.data();
            sa_params p{};
            p.ldmWgt = 1.0;
            p.modVec = mod.data();
            p.modNum = 4;
            p.ynkActPrmVec = ynk.data();
            p.reaActPrmVec = rea.data();
            p.timeNum = 3;
            ids.push_back(srv.add_mesh(m, nullptr, p, lmk.data()));
        }
        // a mesh no session could be built for is refused at registration
        try {
            sa_mesh bad{};
            bad.dim = 3;
            sa_params p{};
            p.modVec = mod.data();
            p.ynkActPrmVec = ynk.data();
            p.reaActPrmVec = rea.data();
            srv.add_mesh(bad, nullptr, p, lmk.data());
            std::fprintf(stderr, "invalid mesh was accepted\n");
            ++failures;
        } catch (const sa::SaError &) {
        }
        // several client threads submit and wait concurrently
        std::vector<std::thread> clients;
        std::atomic<int> bad{0};
        for (int c = 0; c < 3; ++c)
            clients.emplace_back([&, c] {
                std::mt19937 rng(1234 + c);
                std::vector<long long> rids;
                std::vector<std::vector<double>> tvs;
                std::vector<int> ms;
                for (int k = 0; k < nreq / 3; ++k) {
                    const int m = (int)(rng() % ids.size());
                    std::vector<double> tv = {double(rng() % 100) / 10.0, 1.0, 2.0, double(k % 5)};
                    double tp = 1.0;
                    sa_params prm{};
                    prm.ldmWgt = 1.0 + (k % 3);
                    prm.modVec = mod.data();
                    prm.modNum = 4;
                    prm.ynkActPrmVec = ynk.data();
                    prm.reaActPrmVec = rea.data();
                    prm.timeNum = 3;
                    rids.push_back(srv.submit(ids[m], tv.data(), 4, &tp, 1, (k % 2) ? &prm : nullptr));
                    tv.push_back(k % 2 ? prm.ldmWgt : 1.0);
                    tvs.push_back(tv);
                    ms.push_back(m);
                    if (k % 10 == 9) {   // wait for some while others are queued
                        const size_t j = rids.size() - 5;
                        sa::Request r = srv.wait(rids[j]);
                        rids[j] = -1;
                        const bool fits = sizes[ms[j]] * 1000LL <= budget;
                        const double want = stub_obj(tvs[j].data(), 4) * tvs[j][4];
                        if (fits ? (r.rc != SA_OK || r.obj != want) : r.rc != SA_E_ALLOC) bad++;
                    }
                }
                for (size_t j = 0; j < rids.size(); ++j) {
                    if (rids[j] < 0) continue;
                    sa::Request r = srv.wait(rids[j]);
                    const bool fits = sizes[ms[j]] * 1000LL <= budget;
                    const double want = stub_obj(tvs[j].data(), 4) * tvs[j][4];
                    if (fits ? (r.rc != SA_OK || r.obj != want) : r.rc != SA_E_ALLOC) bad++;
                }
            });
        for (auto &t : clients) t.join();
        long long used = 0;
        int sessions = 0, queued = 0, running = 0;
        srv.info(&used, &sessions, &queued, &running);
        if (queued || running || used > budget) ++failures;
        failures += bad.load();
    }
    if (g_sessions.load() != 0 || g_live.load() != 0) {
        std::fprintf(stderr, "sessions left: %d (%lld bytes)\n", g_sessions.load(), g_live.load());
        ++failures;
    }
    failures += g_violations.load();
    std::printf("server stub: %d requests, peak live bytes %lld of budget %lld, failures %d\n", nreq, g_peak.load(),
                budget, failures);
    return failures ? 1 : 0;
}
