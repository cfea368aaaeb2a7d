// Uneven batches of independent transform jobs, split form.
//
// A job is the round trip of nfields fields at one truncation over nlev
// radial levels: synthesis to grids (scal_to_spat, or torpol_to_spat) and
// analysis back (scal_to_SH, or spat_to_qst) -- the calls the reference makes
// per level and field (src/rIter.f90:483-707, src/sht_native.f90).  The batch
// keeps one plan (tables) per distinct l_max and one set of FP32 workspaces
// per job.
//
// Persistent form (the default): one captured launch of k_pbatch runs the
// whole batch as a work list -- synthesis items, row items (inverse and
// forward row FFTs in LDS, so no grid reaches HBM), analysis items -- with
// the tile shapes of every l_max chosen when the batch is made by timing the
// whole launch (autotuner, coordinate descent over the l_max groups):
// orders per synthesis/analysis item (one m, or the pair m, m_max - m whose
// l ranges add up to l_max + 1), level groups of 8 rows per item (1, 2, 4),
// and, where the row FFTs run on the generic LDS passes, theta rows per row
// item (4, 8, 16 as the LDS allows; on wave slots, n_phi = 192, 384, 576,
// 768, one theta row per slot: 32, 16, 8, 8).
//
// Graph form (MHX_BATCH_FORM=graph, or a plan the persistent kernel does not
// admit): every job's four launches (Legendre synthesis, inverse row FFTs to
// grids in HBM, forward row FFTs, Legendre analysis) in one captured graph
// whose jobs lie on parallel branches, the kernel forms chosen per l_max by
// timing the candidates on that l_max's largest job.
#include "mhx_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace mhx {
namespace {

constexpr int NBR = 4;  // parallel branches of the captured graph

constexpr int PB_GROUP = 8;  // rows (field x level) per level group of the persistent form

struct Job {
    mhx_job d;
    int group = 0;
    int ncomp = 1;  // 1 scalar, 3 vector
    int rows = 0;   // nfields * nlev
    float *F = nullptr, *G = nullptr, *F2 = nullptr;
    int *lcut = nullptr;
    int c_syn = 0;  // persistent form: first done counter of the job's level groups
};

struct Group {
    int l_max = 0;
    mhx_plan *plan = nullptr;
    int fft_rows = 4, syn_form = 1, ana_form = 1;  // graph form
    // persistent form tile shapes: orders per item (fold: m with m_max - m),
    // level groups per item (lpi below order index msplit, lpi_hi from it
    // on; lpi divides lpi_hi), theta rows per row item
    int fold = 1, lpi = 1, lpi_hi = 1, msplit = 1 << 30, tbr = 8;
};

template <typename T> void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int grid_n_theta(int l_max) { return 3 * (l_max + 1) / 2; }

bool on_device(const void *ptr)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

}  // namespace
}  // namespace mhx

struct mhx_batch {
    std::vector<mhx::Group> groups;
    std::vector<mhx::Job> jobs;
    hipStream_t cs = nullptr;
    hipStream_t br[mhx::NBR] = {};
    hipEvent_t ev_fork = nullptr, ev_join[mhx::NBR] = {}, ev_in = nullptr, ev_out = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipGraphExec_t exec = nullptr;
    // persistent form
    bool persistent = false;
    mhx::PBJob *d_jobs = nullptr;
    mhx::PBItem *d_items = nullptr;
    size_t items_cap = 0;
    mhx::PBQueues q = {};
    int *d_ctr = nullptr;
    int nctr = 0;
    int pgrid = 0;
};

namespace mhx {
namespace {

// the four launches of one job on stream s, with the group's kernel forms
int launch_job(mhx_batch *b, Job &j, hipStream_t s)
{
    Group &g = b->groups[j.group];
    Plan *P = &g.plan->p;
    const bool vec = j.d.kind == MHX_JOB_VECTOR;
    const size_t lm = (size_t)P->lm_max, R = (size_t)j.rows;
    const int nc = j.ncomp;
    const size_t f_ld = (size_t)nc * P->n_m * P->n_theta;
    const size_t g_ld = (size_t)nc * P->n_phi * P->n_theta, g_comp = (size_t)P->n_phi * P->n_theta;
    const size_t f2_ld = (size_t)nc * P->n_m * ((P->n_theta + 15) / 16 * 16);
    int rc;
    NSynArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int k = 0; k < nc; ++k) a.src[k] = j.d.in + (size_t)k * R * lm * 2;
    a.src_ld = lm;
    a.fout = j.F;
    a.fout_ld = f_ld;
    a.comp0 = 0;
    a.lcut = j.lcut;
    a.or2 = nullptr;
    a.nlev = j.rows;
    P->syn_form = vec ? g.syn_form : 1;  // the scalar alone has the whole-K form only
    if ((rc = launch_nsyn16(P, vec ? NS_QST : NS_S, a, s))) return rc;
    if ((rc = launch_ifft_f32(P, j.F, f_ld / 1, nc, j.rows, j.G, g_ld, g_comp, g.fft_rows, s)))
        return rc;
    if ((rc = launch_fft_f32(P, j.G, g_ld, g_comp, nc, j.rows, vec ? 6u : 0u, j.F2, f2_ld, g.fft_rows, s)))
        return rc;
    NAnaArgs c;
    std::memset(&c, 0, sizeof(c));
    c.fin = j.F2;
    c.fin_ld = f2_ld;
    for (int k = 0; k < 3; ++k) {
        c.in_comp[k] = k < nc ? k : -1;
        c.out[k] = k < nc ? j.d.out + (size_t)k * R * lm * 2 : nullptr;
    }
    c.out_ld = lm;
    c.lcut = j.lcut;
    c.nlev = j.rows;
    P->ana_form = vec ? g.ana_form : 1;
    return launch_nana16(P, vec ? NA_QST : NA_SH, c, s);
}

// the least time of `reps` runs of one job, in ms
int time_job(mhx_batch *b, Job &j, int reps, float &best)
{
    best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        MHX_HIP_CHECK(hipEventRecord(b->t0, b->cs));
        int rc = launch_job(b, j, b->cs);
        if (rc) return rc;
        MHX_HIP_CHECK(hipEventRecord(b->t1, b->cs));
        MHX_HIP_CHECK(hipEventSynchronize(b->t1));
        float ms = 0.f;
        MHX_HIP_CHECK(hipEventElapsedTime(&ms, b->t0, b->t1));
        best = std::min(best, ms);
    }
    return MHX_OK;
}

// per l_max: rows per block of the row FFTs, then (vector jobs) the Legendre
// forms, each chosen by the least time of the group's largest job
int autotune(mhx_batch *b)
{
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        Group &g = b->groups[gi];
        Job *probe = nullptr, *vprobe = nullptr;
        for (Job &j : b->jobs) {
            if (j.group != (int)gi) continue;
            if (!probe || j.rows * j.ncomp > probe->rows * probe->ncomp) probe = &j;
            if (j.d.kind == MHX_JOB_VECTOR && (!vprobe || j.rows > vprobe->rows)) vprobe = &j;
        }
        if (!probe) continue;
        int rc;
        float t, best = 1e30f;
        int best_rows = 1;
        for (int tb : {1, 2, 4, 8}) {
            g.fft_rows = tb;
            if ((rc = time_job(b, *probe, 3, t))) return rc;
            if (t < best) {
                best = t;
                best_rows = tb;
            }
        }
        g.fft_rows = best_rows;
        if (!vprobe) continue;
        float tw, tc;
        g.syn_form = 1;
        g.ana_form = 1;
        if ((rc = time_job(b, *vprobe, 3, tw))) return rc;
        g.syn_form = 2;
        if ((rc = time_job(b, *vprobe, 3, tc))) return rc;
        g.syn_form = tc < tw ? 2 : 1;
        const float tbest = std::min(tw, tc);
        g.ana_form = 2;
        if ((rc = time_job(b, *vprobe, 3, tc))) return rc;
        g.ana_form = tc < tbest ? 2 : 1;
    }
    return MHX_OK;
}

int capture(mhx_batch *b)
{
    hipGraph_t graph = nullptr;
    MHX_HIP_CHECK(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
    int rc = MHX_OK;
    if (hipEventRecord(b->ev_fork, b->cs) != hipSuccess) rc = fail(MHX_E_HIP, "event record failed");
    for (int i = 0; i < NBR && !rc; ++i)
        if (hipStreamWaitEvent(b->br[i], b->ev_fork, 0) != hipSuccess) rc = fail(MHX_E_HIP, "stream wait failed");
    // largest jobs first, dealt round-robin over the branches
    std::vector<int> order(b->jobs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        const Job &p = b->jobs[x], &q = b->jobs[y];
        return (double)p.rows * p.ncomp * p.d.l_max * p.d.l_max > (double)q.rows * q.ncomp * q.d.l_max * q.d.l_max;
    });
    for (size_t i = 0; i < order.size() && !rc; ++i) rc = launch_job(b, b->jobs[order[i]], b->br[i % NBR]);
    for (int i = 0; i < NBR && !rc; ++i) {
        if (hipEventRecord(b->ev_join[i], b->br[i]) != hipSuccess ||
            hipStreamWaitEvent(b->cs, b->ev_join[i], 0) != hipSuccess)
            rc = fail(MHX_E_HIP, "graph join failed");
    }
    const hipError_t ec = hipStreamEndCapture(b->cs, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(MHX_E_HIP, "graph capture of the batch failed");
    const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(MHX_E_HIP, "graph instantiation of the batch failed");
    return MHX_OK;
}

// ------------------------------------------------------- persistent form
int ngroups(const Job &j) { return (j.rows + PB_GROUP - 1) / PB_GROUP; }

// theta rows per row item that the launch's LDS holds at this l_max
bool tbr_fits(const Plan *P, int tbr) { return pbatch_row_lds(P->M, tbr) <= pbatch_lds(); }

// the jobs' descriptors with the groups' current row-item shapes
void make_pjobs(const mhx_batch *b, std::vector<PBJob> &out)
{
    out.assign(b->jobs.size(), PBJob());
    for (size_t ji = 0; ji < b->jobs.size(); ++ji) {
        const Job &j = b->jobs[ji];
        const Group &g = b->groups[j.group];
        const Plan *P = &g.plan->p;
        PBJob &pj = out[ji];
        std::memset(&pj, 0, sizeof(pj));
        const bool vec = j.d.kind == MHX_JOB_VECTOR;
        const size_t lm = (size_t)P->lm_max, R = (size_t)j.rows;
        const int nc = j.ncomp;
        for (int k = 0; k < nc; ++k) pj.syn.src[k] = j.d.in + (size_t)k * R * lm * 2;
        pj.syn.src_ld = lm;
        pj.syn.fout = j.F;
        pj.syn.fout_ld = (size_t)nc * P->n_m * P->n_theta;
        pj.syn.lcut = j.lcut;
        pj.syn.nlev = j.rows;
        pj.ana.fin = j.F2;
        pj.ana.fin_ld = (size_t)nc * P->n_m * ((P->n_theta + 15) / 16 * 16);
        for (int k = 0; k < 3; ++k) {
            pj.ana.in_comp[k] = k < nc ? k : -1;
            pj.ana.out[k] = k < nc ? j.d.out + (size_t)k * R * lm * 2 : nullptr;
        }
        pj.ana.out_ld = lm;
        pj.ana.lcut = j.lcut;
        pj.ana.nlev = j.rows;
        (void)fill_pbatch_job(P, pj);  // admitted when the form was chosen
        pj.syn_kind = vec ? NS_QST : NS_S;
        pj.ana_kind = vec ? NA_QST : NA_SH;
        pj.osin2_mask = vec ? 6u : 0u;
        pj.F = j.F;
        pj.F2 = j.F2;
        pj.f_ld = pj.syn.fout_ld;
        pj.f2_ld = pj.ana.fin_ld;
        pj.nc = nc;
        pj.rows = j.rows;
        pj.tbr = g.tbr;
        pj.ntbr = (P->n_theta + g.tbr - 1) / g.tbr;
        pj.c_syn = j.c_syn;
        pj.c_row = j.c_syn + ngroups(j);
    }
}

// The work list with the groups' current tile shapes.  A chain is the
// synthesis, row and analysis items of one job's run of lpi_hi level groups
// (its orders below msplit in items of lpi groups, the others in items of
// the whole run); chains are dealt to the XCD queues largest first, each to
// the least loaded queue (estimated work: l range x rows x components for
// the Legendre items, rows x components x n_phi for the row items); within
// a queue the synthesis items come first, then the row items, then the
// analysis items, each phase largest first.
void make_items(const mhx_batch *b, std::vector<PBItem> &items, PBQueues &qs)
{
    struct W {
        double cost;
        PBItem it;
    };
    struct Chain {
        double cost = 0.0;
        std::vector<W> syn, row, ana;
    };
    std::vector<Chain> chains;
    for (size_t ji = 0; ji < b->jobs.size(); ++ji) {
        const Job &j = b->jobs[ji];
        const Group &g = b->groups[j.group];
        const Plan *P = &g.plan->p;
        const int ng = ngroups(j), n_m = P->n_m;
        const double wc = j.ncomp == 3 ? 3.0 : 1.0;
        std::vector<std::pair<int, int>> ords;
        if (g.fold) {
            for (int mc = 0; mc < n_m / 2; ++mc) ords.push_back({mc, n_m - 1 - mc});
            if (n_m % 2) ords.push_back({n_m / 2, -1});
        } else {
            for (int mc = 0; mc < n_m; ++mc) ords.push_back({mc, -1});
        }
        const int ntbr = (P->n_theta + g.tbr - 1) / g.tbr;
        const int span = std::max(g.lpi, g.lpi_hi);
        auto rows_in = [&](int g0, int gn) {
            int nr = 0;
            for (int gg = g0; gg < g0 + gn; ++gg) nr += std::min(PB_GROUP, j.rows - gg * PB_GROUP);
            return nr;
        };
        for (int c0 = 0; c0 < ng; c0 += span) {
            Chain ch;
            const int cn = std::min(span, ng - c0);
            for (const auto &o : ords) {
                const int step = o.first >= g.msplit ? g.lpi_hi : g.lpi;
                for (int g0 = c0; g0 < c0 + cn; g0 += step) {
                    const int gn = std::min(step, c0 + cn - g0);
                    PBItem it;
                    std::memset(&it, 0, sizeof(it));
                    it.kind = PB_SYN;
                    it.job = (short)ji;
                    it.m0 = (short)o.first;
                    it.m1 = (short)o.second;
                    it.g0 = (short)g0;
                    it.ng = (short)gn;
                    double kl = (double)(P->l_max + 2 - o.first * P->minc);
                    if (o.second >= 0) kl += (double)(P->l_max + 2 - o.second * P->minc);
                    const double c = kl * rows_in(g0, gn) * wc * P->n_theta;
                    ch.syn.push_back({c, it});
                    it.kind = PB_ANA;
                    ch.ana.push_back({c, it});
                    ch.cost += 2 * c;
                }
            }
            const int g0 = c0, gn = cn;
            for (int gg = g0; gg < g0 + gn; ++gg) {
                const int nrg = std::min(PB_GROUP, j.rows - gg * PB_GROUP);
                for (int tb = 0; tb < ntbr; ++tb) {
                    PBItem it;
                    std::memset(&it, 0, sizeof(it));
                    it.kind = PB_ROW;
                    it.job = (short)ji;
                    it.m0 = it.m1 = -1;
                    it.g0 = (short)gg;
                    it.ng = 1;
                    it.tb = (short)tb;
                    const double c = (double)nrg * j.ncomp * P->n_phi * g.tbr;
                    ch.row.push_back({c, it});
                    ch.cost += c;
                }
            }
            chains.push_back(std::move(ch));
        }
    }
    std::vector<int> order(chains.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return chains[x].cost > chains[y].cost; });
    const int nq = std::max(1, b->q.nq);
    std::vector<double> load(nq, 0.0);
    std::vector<std::vector<int>> mine(nq);
    for (int ci : order) {
        const int qi = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[qi] += chains[ci].cost;
        mine[qi].push_back(ci);
    }
    auto by_cost = [](const W &x, const W &y) { return x.cost > y.cost; };
    items.clear();
    qs = b->q;
    for (int qi = 0; qi < nq; ++qi) {
        qs.off[qi] = (int)items.size();
        std::vector<W> syn, row, ana;
        for (int ci : mine[qi]) {
            syn.insert(syn.end(), chains[ci].syn.begin(), chains[ci].syn.end());
            row.insert(row.end(), chains[ci].row.begin(), chains[ci].row.end());
            ana.insert(ana.end(), chains[ci].ana.begin(), chains[ci].ana.end());
        }
        std::stable_sort(syn.begin(), syn.end(), by_cost);
        std::stable_sort(row.begin(), row.end(), by_cost);
        std::stable_sort(ana.begin(), ana.end(), by_cost);
        for (auto *v : {&syn, &row, &ana})
            for (const W &w : *v) items.push_back(w.it);
    }
    qs.off[nq] = (int)items.size();
}

// uploads the descriptors and the work list of the current shapes
int stage(mhx_batch *b)
{
    std::vector<PBJob> pj;
    std::vector<PBItem> items;
    make_pjobs(b, pj);
    make_items(b, items, b->q);
    if (items.size() > b->items_cap) {
        dfree(b->d_items);
        if (hipMalloc(&b->d_items, items.size() * sizeof(PBItem)) != hipSuccess)
            return fail(MHX_E_NOMEM, "batch work-list allocation failed");
        b->items_cap = items.size();
    }
    if (!b->d_jobs && hipMalloc(&b->d_jobs, pj.size() * sizeof(PBJob)) != hipSuccess)
        return fail(MHX_E_NOMEM, "batch descriptor allocation failed");
    if (hipMemcpy(b->d_jobs, pj.data(), pj.size() * sizeof(PBJob), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_items, items.data(), items.size() * sizeof(PBItem), hipMemcpyHostToDevice) != hipSuccess)
        return fail(MHX_E_HIP, "batch work-list upload failed");
    return MHX_OK;
}

int launch_p(mhx_batch *b, hipStream_t s)
{
    return launch_pbatch(b->d_jobs, b->d_items, b->q, b->d_ctr, b->nctr, b->pgrid, s);
}

// MHX_OK, or an error if a wait of the persistent launch ran out (the flag is
// cleared)
int check_flag(mhx_batch *b)
{
    int flag = 0;
    MHX_HIP_CHECK(hipStreamSynchronize(b->cs));
    MHX_HIP_CHECK(hipMemcpy(&flag, b->d_ctr + PB_ERR, sizeof(int), hipMemcpyDeviceToHost));
    if (!flag) return MHX_OK;
    MHX_HIP_CHECK(hipMemset(b->d_ctr + PB_ERR, 0, sizeof(int)));
    return fail(MHX_E_HIP, (flag & 1) ? "a work item of the persistent batch launch gave up waiting for its "
                                        "inputs; the outputs of that execution are not valid"
                                      : "a work queue of the persistent batch launch was left unclaimed (an "
                                        "XCD ran none of its workgroups); the outputs are not valid");
}

// the least time of `reps` launches of the staged work list, in ms
int time_p(mhx_batch *b, int reps, float &best)
{
    best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        MHX_HIP_CHECK(hipEventRecord(b->t0, b->cs));
        int rc = launch_p(b, b->cs);
        if (rc) return rc;
        MHX_HIP_CHECK(hipEventRecord(b->t1, b->cs));
        MHX_HIP_CHECK(hipEventSynchronize(b->t1));
        float ms = 0.f;
        MHX_HIP_CHECK(hipEventElapsedTime(&ms, b->t0, b->t1));
        best = std::min(best, ms);
    }
    return check_flag(b);
}

// tile shapes per l_max group by coordinate descent on the whole launch:
// the Legendre items' shape (orders per item, level groups per item below
// and from the split order), then theta rows per row item
int autotune_p(mhx_batch *b)
{
    for (Group &g : b->groups) {
        g.fold = 1;
        g.lpi = g.lpi_hi = 1;
        g.msplit = 1 << 30;
        const int ns = pbatch_slot_rows(g.plan->p.M);
        g.tbr = ns ? ns : (tbr_fits(&g.plan->p, 8) ? 8 : 4);
    }
    int rc;
    float best;
    if ((rc = stage(b)) || (rc = time_p(b, 2, best))) return rc;
    // {fold, lpi, lpi_hi, split at n_m / d (0: none)}: folded pairs of even
    // cost, or single orders whose short high-m ones share items over more
    // level groups
    const int shapes[7][4] = {{1, 1, 1, 0}, {1, 2, 2, 0}, {1, 4, 4, 0}, {0, 1, 1, 0},
                              {0, 1, 2, 2}, {0, 1, 4, 2}, {0, 1, 4, 4}};
    for (Group &g : b->groups) {
        const int n_m = g.plan->p.n_m;
        auto set = [&](const int *sh) {
            g.fold = sh[0];
            g.lpi = sh[1];
            g.lpi_hi = sh[2];
            g.msplit = sh[3] ? n_m / sh[3] : 1 << 30;
        };
        int bi = 0;  // the start shape is shapes[0]
        for (int si = 1; si < 7; ++si) {
            set(shapes[si]);
            float t;
            if ((rc = stage(b)) || (rc = time_p(b, 2, t))) return rc;
            if (t < best) {
                best = t;
                bi = si;
            }
        }
        set(shapes[bi]);
        if (pbatch_slot_rows(g.plan->p.M)) continue;  // slot row items: theta rows fixed
        const int t0 = g.tbr;
        int bt = t0;
        for (int tbr : {4, 8, 16}) {
            if (tbr == t0 || !tbr_fits(&g.plan->p, tbr)) continue;
            g.tbr = tbr;
            float t;
            if ((rc = stage(b)) || (rc = time_p(b, 2, t))) return rc;
            if (t < best) {
                best = t;
                bt = tbr;
            }
        }
        g.tbr = bt;
    }
    return stage(b);
}

int capture_p(mhx_batch *b)
{
    hipGraph_t graph = nullptr;
    MHX_HIP_CHECK(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
    const int rc = launch_p(b, b->cs);
    const hipError_t ec = hipStreamEndCapture(b->cs, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(MHX_E_HIP, "graph capture of the batch failed");
    const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(MHX_E_HIP, "graph instantiation of the batch failed");
    return MHX_OK;
}

}  // namespace
}  // namespace mhx

using namespace mhx;

extern "C" {

int mhx_batch_create(const mhx_job *jobs, int njobs, int precision, mhx_batch **batch)
{
    if (!batch) return fail(MHX_E_INVALID, "null batch pointer");
    *batch = nullptr;
    if (!jobs || njobs <= 0) return fail(MHX_E_INVALID, "no jobs");
    if (precision != MHX_PREC_SPLIT16)
        return fail(MHX_E_UNSUPPORTED, "batches run in the split form (MHX_PREC_SPLIT16)");
    for (int i = 0; i < njobs; ++i) {
        const mhx_job &d = jobs[i];
        if (d.l_max < 1 || d.l_max > 255 || (d.l_max + 1) % 2)
            return fail(MHX_E_INVALID, "job " + std::to_string(i) + ": l_max must be odd, 1 .. 255");
        if (d.nlev <= 0 || d.nfields <= 0 || !d.in || !d.out)
            return fail(MHX_E_INVALID, "job " + std::to_string(i) + ": bad sizes or null arrays");
        if (d.kind != MHX_JOB_SCALAR && d.kind != MHX_JOB_VECTOR)
            return fail(MHX_E_INVALID, "job " + std::to_string(i) + ": unknown kind");
        if (!on_device(d.in) || !on_device(d.out))
            return fail(MHX_E_INVALID, "job " + std::to_string(i) + ": arrays must be device memory");
    }
    mhx_batch *b = new mhx_batch();
    auto bail = [&](int rc) {
        mhx_batch_destroy(b);
        return rc;
    };
    int rc;
    for (int i = 0; i < njobs; ++i) {
        Job j;
        j.d = jobs[i];
        j.ncomp = j.d.kind == MHX_JOB_VECTOR ? 3 : 1;
        j.rows = j.d.nfields * j.d.nlev;
        int gi = -1;
        for (size_t k = 0; k < b->groups.size(); ++k)
            if (b->groups[k].l_max == j.d.l_max) gi = (int)k;
        if (gi < 0) {
            Group g;
            g.l_max = j.d.l_max;
            const int nt = grid_n_theta(g.l_max);
            if ((rc = mhx_plan_create(g.l_max, g.l_max, 1, nt, 2 * nt, MHX_PREC_SPLIT16, 1, &g.plan))) {
                b->groups.push_back(g);
                return bail(rc);
            }
            b->groups.push_back(g);
            gi = (int)b->groups.size() - 1;
        }
        j.group = gi;
        b->jobs.push_back(j);
    }
    // the persistent form unless asked for the graph form or a plan does not
    // admit it
    {
        const char *e = std::getenv("MHX_BATCH_FORM");
        b->persistent = !(e && std::strcmp(e, "graph") == 0);
        for (const Group &g : b->groups) b->persistent = b->persistent && pbatch_supported(&g.plan->p);
    }
    for (Job &j : b->jobs) {
        const Plan *P = &b->groups[j.group].plan->p;
        const size_t R = (size_t)j.rows, nc = (size_t)j.ncomp;
        const size_t fF = R * nc * P->n_m * P->n_theta * 2;
        const size_t fG = R * nc * P->n_phi * P->n_theta;
        const size_t fF2 = R * nc * P->n_m * ((P->n_theta + 15) / 16 * 16) * 2;
        if (hipMalloc(&j.F, fF * sizeof(float)) != hipSuccess ||
            (!b->persistent && hipMalloc(&j.G, fG * sizeof(float)) != hipSuccess) ||
            hipMalloc(&j.F2, fF2 * sizeof(float)) != hipSuccess ||
            hipMalloc(&j.lcut, R * sizeof(int)) != hipSuccess)
            return bail(fail(MHX_E_NOMEM, "batch workspace allocation failed"));
        std::vector<int> lc(R, j.d.l_max);
        if (hipMemcpy(j.lcut, lc.data(), R * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(MHX_E_HIP, "lcut upload failed"));
    }
    if (hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(MHX_E_HIP, "stream creation failed"));
    for (int i = 0; i < NBR; ++i)
        if (hipStreamCreateWithFlags(&b->br[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming) != hipSuccess)
            return bail(fail(MHX_E_HIP, "stream or event creation failed"));
    if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess)
        return bail(fail(MHX_E_HIP, "event creation failed"));
    if (b->persistent) {
        b->nctr = PB_CTR0;
        for (Job &j : b->jobs) {
            j.c_syn = b->nctr;
            b->nctr += 2 * ngroups(j);
        }
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            return bail(fail(MHX_E_HIP, "device query failed"));
        b->pgrid = 2 * ncu;  // two resident workgroups per CU
        if (hipMalloc(&b->d_ctr, (size_t)std::max(b->nctr, b->pgrid) * sizeof(int)) != hipSuccess)
            return bail(fail(MHX_E_NOMEM, "batch counter allocation failed"));
        // the XCDs the launch's workgroups run on: one queue per XCC_ID seen
        {
            std::vector<int> ids(b->pgrid);
            if ((rc = probe_xcc_ids(b->d_ctr, b->pgrid, b->cs))) return bail(rc);
            if (hipMemcpyAsync(ids.data(), b->d_ctr, ids.size() * sizeof(int), hipMemcpyDeviceToHost, b->cs) !=
                    hipSuccess ||
                hipStreamSynchronize(b->cs) != hipSuccess)
                return bail(fail(MHX_E_HIP, "XCD probe failed"));
            for (int &x : b->q.xmap) x = -1;
            b->q.nq = 0;
            for (int id : ids) {
                const int x = id & (PB_MAXQ - 1);
                if (b->q.xmap[x] < 0) b->q.xmap[x] = b->q.nq++;
            }
        }
        if (hipMemset(b->d_ctr, 0, (size_t)b->nctr * sizeof(int)) != hipSuccess)
            return bail(fail(MHX_E_HIP, "batch counter clear failed"));
        if ((rc = autotune_p(b)) || (rc = capture_p(b))) return bail(rc);
    } else if ((rc = autotune(b)) || (rc = capture(b))) {
        return bail(rc);
    }
    *batch = b;
    return MHX_OK;
}

int mhx_batch_execute(mhx_batch *b, void *stream)
{
    if (!b || !b->exec) return fail(MHX_E_INVALID, "null batch");
    if (stream) {
        MHX_HIP_CHECK(hipEventRecord(b->ev_in, (hipStream_t)stream));
        MHX_HIP_CHECK(hipStreamWaitEvent(b->cs, b->ev_in, 0));
    }
    MHX_HIP_CHECK(hipGraphLaunch(b->exec, b->cs));
    if (stream) {
        MHX_HIP_CHECK(hipEventRecord(b->ev_out, b->cs));
        MHX_HIP_CHECK(hipStreamWaitEvent((hipStream_t)stream, b->ev_out, 0));
        return MHX_OK;
    }
    MHX_HIP_CHECK(hipStreamSynchronize(b->cs));
    return MHX_OK;
}

int mhx_batch_tuning(const mhx_batch *b, int *l_max, int *syn_form, int *ana_form, int *fft_rows,
                     int n)
{
    if (!b) return fail(MHX_E_INVALID, "null batch");
    const int k = std::min(n, (int)b->groups.size());
    for (int i = 0; i < k; ++i) {
        const Group &g = b->groups[i];
        if (l_max) l_max[i] = g.l_max;
        if (syn_form) syn_form[i] = g.syn_form;
        if (ana_form) ana_form[i] = g.ana_form;
        if (fft_rows) fft_rows[i] = b->persistent ? g.tbr : g.fft_rows;
    }
    return (int)b->groups.size();
}

int mhx_batch_shapes(const mhx_batch *b, int *form, int *l_max, int *orders, int *groups,
                     int *groups_high, int *m_split, int *rows, int n)
{
    if (!b) return fail(MHX_E_INVALID, "null batch");
    if (form) *form = b->persistent ? 1 : 0;
    const int k = std::min(n, (int)b->groups.size());
    for (int i = 0; i < k; ++i) {
        const Group &g = b->groups[i];
        if (l_max) l_max[i] = g.l_max;
        if (orders) orders[i] = b->persistent ? (g.fold ? 2 : 1) : 1;
        if (groups) groups[i] = b->persistent ? g.lpi : 0;
        if (groups_high) groups_high[i] = b->persistent ? g.lpi_hi : 0;
        if (m_split) m_split[i] = b->persistent ? std::min(g.msplit, g.plan->p.n_m) : 0;
        if (rows) rows[i] = b->persistent ? g.tbr : g.fft_rows;
    }
    return (int)b->groups.size();
}

int mhx_batch_status(mhx_batch *b)
{
    if (!b) return fail(MHX_E_INVALID, "null batch");
    if (!b->persistent) {
        MHX_HIP_CHECK(hipStreamSynchronize(b->cs));
        return MHX_OK;
    }
    return check_flag(b);
}

int mhx_batch_destroy(mhx_batch *b)
{
    if (!b) return MHX_OK;
    if (b->cs) (void)hipStreamSynchronize(b->cs);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    for (Job &j : b->jobs) {
        dfree(j.F);
        dfree(j.G);
        dfree(j.F2);
        dfree(j.lcut);
    }
    dfree(b->d_jobs);
    dfree(b->d_items);
    dfree(b->d_ctr);
    for (Group &g : b->groups)
        if (g.plan) mhx_plan_destroy(g.plan);
    for (int i = 0; i < NBR; ++i) {
        if (b->ev_join[i]) (void)hipEventDestroy(b->ev_join[i]);
        if (b->br[i]) (void)hipStreamDestroy(b->br[i]);
    }
    for (hipEvent_t e : {b->ev_fork, b->ev_in, b->ev_out, b->t0, b->t1})
        if (e) (void)hipEventDestroy(e);
    if (b->cs) (void)hipStreamDestroy(b->cs);
    delete b;
    return MHX_OK;
}

}  // extern "C"
