// Uneven batches of independent transform jobs, split form.
//
// A job is the round trip of nfields fields at one truncation over nlev
// radial levels: synthesis to grids (scal_to_spat, or torpol_to_spat) and
// analysis back (scal_to_SH, or spat_to_qst) -- the calls the reference makes
// per level and field (src/rIter.f90:483-707, src/sht_native.f90).  The batch
// keeps one plan (tables) per distinct l_max, one set of FP32 workspaces per
// job, and runs every job's four launches (Legendre synthesis, inverse row
// FFTs, forward row FFTs, Legendre analysis) from one captured graph whose
// jobs lie on parallel branches, so the small jobs of low l_max fill the GPU
// beside the large ones.  The kernel forms are chosen per l_max when the
// batch is made, by timing the candidates on that l_max's largest job
// (autotuner): rows per block of the row FFTs, and for vector jobs the
// whole-K/whole-theta or the chunked Legendre kernels.
#include "mhx_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace mhx {
namespace {

constexpr int NBR = 4;  // parallel branches of the captured graph

struct Job {
    mhx_job d;
    int group = 0;
    int ncomp = 1;  // 1 scalar, 3 vector
    int rows = 0;   // nfields * nlev
    float *F = nullptr, *G = nullptr, *F2 = nullptr;
    int *lcut = nullptr;
};

struct Group {
    int l_max = 0;
    mhx_plan *plan = nullptr;
    int fft_rows = 4, syn_form = 1, ana_form = 1;
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
    for (Job &j : b->jobs) {
        const Plan *P = &b->groups[j.group].plan->p;
        const size_t R = (size_t)j.rows, nc = (size_t)j.ncomp;
        const size_t fF = R * nc * P->n_m * P->n_theta * 2;
        const size_t fG = R * nc * P->n_phi * P->n_theta;
        const size_t fF2 = R * nc * P->n_m * ((P->n_theta + 15) / 16 * 16) * 2;
        if (hipMalloc(&j.F, fF * sizeof(float)) != hipSuccess ||
            hipMalloc(&j.G, fG * sizeof(float)) != hipSuccess ||
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
    if ((rc = autotune(b)) || (rc = capture(b))) return bail(rc);
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
        if (fft_rows) fft_rows[i] = g.fft_rows;
    }
    return (int)b->groups.size();
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
