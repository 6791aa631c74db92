// Batched MLEM over many independent problems of uneven size (declared in
// include/fl_correction_amd.h as fl_mlem_multi_*).
//
// A problem is one emission line of one sample (R = 1): atten (A,B,H,W),
// em (A), theta (A), measured I (A,B,W), estimate C (B,H,W).  The reference
// reconstructs them one call at a time (mlem_cuda_batch,
// FL/cuda_lib/atten_cuda.cpp:283-345, behind torch_mlem_recon_batch,
// FL/torch_recon.py:50-98); here every pass of an iteration is ONE persistent
// launch over all problems:
//
//   ratio pass:  ratio_i = I_i / max(F_i C_i, 1e-6)          grouped forward
//   update pass: C'_i = max(0, C_i * B_i^T ratio_i / (sens_i + beta*huber(C_i)))
//                                                             grouped back-projector
//
// The plan builds the job tables once on the host and uploads them into the
// caller's workspace: five tables of launch records (forward even/odd, back-
// projector sensitivity/even/odd; even reads C and writes the workspace copy,
// odd the reverse) plus the per-job tile prefix sums.  Two iterations (four
// kernel launches, no memset or copy nodes) are captured once into a hipGraph
// and replayed.  No device memory is allocated: per-problem sensitivity,
// ratio and the second estimate live in the workspace.
//
// Slice-major mode (taken when the workspace holds
// fl_mlem_multi_workspace_bytes_sm): problems with the same theta array and
// the same A, H, W -- the emission lines of one sample -- are packed along the
// slice axis into one problem of sum(B) slices, whose slice-major attenuation
// copy (A, H, W, sum B) lives in the workspace.  Slices are independent in
// both passes, so a pack is exactly its problems side by side.  The forward
// writes em[a] * ratio into the pack's ratio (A, W, sum B) at the problem's
// slice offset, and the packs are back-projected by the grouped slice-major
// kernels with em = 1; the sensitivity is the back-projection of a ratio
// filled with each problem's em.  The estimates iterate in two packed copies
// in the workspace: a run copies C in at the start and the result out at the
// end (the copies stay outside the captured graph).
#include "projector.hpp"

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

struct fl_mlem_multi {
    int device = 0;
    int n = 0;
    fl::FwdArgs *fwd[2] = {nullptr, nullptr};
    fl::BwdArgs *bwd[2] = {nullptr, nullptr};
    fl::BwdArgs *sens = nullptr;
    fl::CopyJob *copy = nullptr;
    uint32_t *fwd_begin = nullptr, *bwd_begin = nullptr, *copy_begin = nullptr;
    uint32_t fwd_tiles = 0, bwd_tiles = 0, copy_tiles = 0;
    int fwd_grid = 1, bwd_grid = 1, copy_grid = 1;
    // slice-major mode
    bool sm = false;
    fl::SmGroup smg;
    fl::CopyJob *pack_in = nullptr, *unpack[2] = {nullptr, nullptr};
    std::vector<char> staging;  // host image of the tables: an async upload reads it
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace fl {
namespace {

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

int fwd_angles_of(const fl_mlem_problem &p) { return p.fwd_angles == 0 ? kFG : p.fwd_angles; }
int bwd_rows_of(const fl_mlem_problem &p) { return p.bwd_rows == 0 ? kBI : p.bwd_rows; }

// Problems packed along the slice axis (slice-major mode): same theta array and
// the same A, H, W; members in input order, member m at slice offset b0[m].
// The pack's slices are padded to Bp, a multiple of its slices per lane.
struct Pack {
    std::vector<int64_t> members, b0;
    int64_t A = 0, H = 0, W = 0, B = 0, Bp = 0;
    int spl = 1;
    const float *theta = nullptr;
};

// Slices per lane of a pack of B slices: the fewest wavefronts per block
// (64 * spl slices each), then the fewest slices per lane (fewer registers,
// more wavefronts resident).  A wavefront's cost hardly depends on how many
// of its lanes hold slices.
int pack_spl(int64_t B)
{
    int best = 1;
    int64_t groups = (B + 63) / 64;
    for (int spl : {2, 4}) {
        const int64_t g = (B + 64 * spl - 1) / (64 * spl);
        if (g < groups) {
            groups = g;
            best = spl;
        }
    }
    return best;
}

std::vector<Pack> make_packs(const fl_mlem_problem *pr, int64_t n)
{
    std::vector<Pack> packs;
    for (int64_t i = 0; i < n; ++i) {
        const fl_mlem_problem &q = pr[i];
        Pack *pk = nullptr;
        for (Pack &c : packs)
            if (c.theta == q.theta && c.A == q.A && c.H == q.H && c.W == q.W) pk = &c;
        if (!pk) {
            packs.emplace_back();
            pk = &packs.back();
            pk->A = q.A;
            pk->H = q.H;
            pk->W = q.W;
            pk->theta = q.theta;
        }
        pk->members.push_back(i);
        pk->b0.push_back(pk->B);
        pk->B += q.B;
    }
    for (Pack &pk : packs) {
        pk.spl = pack_spl(pk.B);
        pk.Bp = (pk.B + pk.spl - 1) / pk.spl * pk.spl;
    }
    return packs;
}

// Byte offsets of the workspace regions.
struct Layout {
    size_t fwd[2], bwd[2], sens_jobs, copy, fwd_begin, bwd_begin, copy_begin;
    // slice-major mode: grouped back-projection tables, pack-in / unpack copy jobs, em = 1
    size_t sm_tables = 0, pack_in = 0, unpack[2] = {0, 0}, ones = 0;
    size_t tables;
    std::vector<size_t> sens, alt, ratio;  // per problem (per pack in slice-major mode)
    std::vector<size_t> atten_sm, alt2, scratch;  // per pack, slice-major mode
    size_t total;
    bool ok = true;  // false: the slice-major mode does not apply
};

Layout layout(const fl_mlem_problem *pr, int64_t n, bool sm)
{
    Layout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += align256(bytes);
        return at;
    };
    L.fwd[0] = take(n * sizeof(FwdArgs));
    L.fwd[1] = take(n * sizeof(FwdArgs));
    L.sens_jobs = take(n * sizeof(BwdArgs));
    L.bwd[0] = take(n * sizeof(BwdArgs));
    L.bwd[1] = take(n * sizeof(BwdArgs));
    L.copy = take(n * sizeof(CopyJob));
    L.fwd_begin = take((n + 1) * sizeof(uint32_t));
    L.bwd_begin = take((n + 1) * sizeof(uint32_t));
    L.copy_begin = take((n + 1) * sizeof(uint32_t));
    if (!sm) {
        L.tables = o;
        for (int64_t i = 0; i < n; ++i) {
            const size_t nC = (size_t)(pr[i].B * pr[i].H * pr[i].W) * sizeof(float);
            const size_t nP = (size_t)(pr[i].A * pr[i].B * pr[i].W) * sizeof(float);
            L.sens.push_back(take(nC));
            L.alt.push_back(take(nC));
            L.ratio.push_back(take(nP));
        }
        L.total = o;
        return L;
    }
    const std::vector<Pack> packs = make_packs(pr, n);
    int64_t max_a = 1;
    for (const Pack &pk : packs) {
        max_a = std::max(max_a, pk.A);
        if (pk.A >= 65536 || pk.H * pk.W * pk.Bp >= ((int64_t)1 << 29)) L.ok = false;
    }
    L.sm_tables = take(sm_group_table_bytes((int)packs.size()));
    L.pack_in = take(n * sizeof(CopyJob));
    L.unpack[0] = take(n * sizeof(CopyJob));
    L.unpack[1] = take(n * sizeof(CopyJob));
    L.ones = take((size_t)max_a * sizeof(float));
    L.tables = o;
    for (const Pack &pk : packs) {
        const size_t nC = (size_t)(pk.Bp * pk.H * pk.W) * sizeof(float);
        L.atten_sm.push_back(take((size_t)(pk.A * pk.H * pk.W * pk.Bp) * sizeof(float)));
        L.ratio.push_back(take((size_t)(pk.A * pk.W * pk.Bp) * sizeof(float)));
        L.sens.push_back(take(nC));
        L.alt.push_back(take(nC));
        L.alt2.push_back(take(nC));
        L.scratch.push_back(L.ok ? take(sm_scratch_bytes(pk.A, pk.Bp, pk.H, pk.W, pk.spl)) : 0);
    }
    L.total = o;
    return L;
}

int check_problems(const fl_mlem_problem *pr, int64_t n)
{
    FL_REQUIRE(pr != nullptr && n >= 1 && n < (1 << 20), "batch: need 1 <= n < 2^20 problems");
    for (int64_t i = 0; i < n; ++i) {
        const fl_mlem_problem &p = pr[i];
        const std::string at = "batch problem " + std::to_string(i) + ": ";
        if (p.A < 1 || p.B < 1 || p.H < 2 || p.W < 2 || p.H >= (1 << 20) || p.W >= (1 << 20) ||
            p.A >= (1 << 24) || p.B >= (1 << 24) || p.H * p.W >= ((int64_t)1 << 29)) {
            set_error(at + "shape: need A >= 1, B >= 1, H >= 2, W >= 2, H*W < 2^29");
            return FL_ERR_INVALID_ARG;
        }
        if (!p.atten || !p.em_cs || !p.theta || !p.I || !p.C) {
            set_error(at + "null pointer argument");
            return FL_ERR_INVALID_ARG;
        }
        if (!(p.fwd_angles == 0 || p.fwd_angles == 4 || p.fwd_angles == 8) ||
            !(p.bwd_rows == 0 || p.bwd_rows == 15 || p.bwd_rows == 31)) {
            set_error(at + "tile shape: fwd_angles must be 0, 4 or 8 and bwd_rows 0, 15 or 31");
            return FL_ERR_INVALID_ARG;
        }
    }
    // Estimates are written: their ranges must be disjoint.
    std::vector<std::pair<uintptr_t, uintptr_t>> r;
    for (int64_t i = 0; i < n; ++i) {
        const uintptr_t b = (uintptr_t)pr[i].C;
        r.emplace_back(b, b + (uintptr_t)(pr[i].B * pr[i].H * pr[i].W) * sizeof(float));
    }
    std::sort(r.begin(), r.end());
    for (size_t i = 1; i < r.size(); ++i)
        FL_REQUIRE(r[i].first >= r[i - 1].second, "batch: the estimates C of two problems overlap");
    return FL_OK;
}

// Jobs ordered by cost per tile, largest first (stable: ties keep input order).
std::vector<int64_t> order_by(int64_t n, const std::vector<double> &cost)
{
    std::vector<int64_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return cost[a] > cost[b]; });
    return ord;
}

int enqueue_iteration(fl_mlem_multi *m, int parity, hipStream_t s)
{
    if (int e = launch_fwd_group(m->fwd[parity], m->fwd_begin, m->n, m->fwd_tiles, m->fwd_grid, s)) return e;
    if (m->sm) return launch_sm_group(m->smg, 1 + parity, s);
    return launch_bwd_group(m->bwd[parity], m->bwd_begin, m->n, m->bwd_tiles, m->bwd_grid, s);
}

}  // namespace
}  // namespace fl

using namespace fl;

extern "C" {

static bool dims_nonnegative(const fl_mlem_problem *problems, int64_t n)
{
    if (!problems || n < 1 || n >= (1 << 20)) return false;
    for (int64_t i = 0; i < n; ++i)
        if (problems[i].A < 0 || problems[i].B < 0 || problems[i].H < 0 || problems[i].W < 0) return false;
    return true;
}

size_t fl_mlem_multi_workspace_bytes(const fl_mlem_problem *problems, int64_t n)
{
    return dims_nonnegative(problems, n) ? layout(problems, n, false).total : 0;
}

size_t fl_mlem_multi_workspace_bytes_sm(const fl_mlem_problem *problems, int64_t n)
{
    if (!dims_nonnegative(problems, n)) return 0;
    const Layout L = layout(problems, n, true);
    return L.ok ? L.total : 0;
}

int fl_mlem_multi_create(const fl_mlem_problem *problems, int64_t n, double beta, double delta, void *workspace,
                         size_t workspace_bytes, void *stream, fl_mlem_multi **plan)
{
    set_error("");
    FL_REQUIRE(plan != nullptr, "null plan pointer");
    *plan = nullptr;
    if (int e = check_problems(problems, n)) return e;
    const Layout Ls = layout(problems, n, true);
    const bool sm = Ls.ok && workspace_bytes >= Ls.total;
    const Layout L = sm ? Ls : layout(problems, n, false);
    if (!workspace || workspace_bytes < L.total) {
        set_error("workspace too small: need " + std::to_string(L.total) + " bytes");
        return FL_ERR_WORKSPACE;
    }
    FL_REQUIRE(((uintptr_t)workspace & 255) == 0, "workspace must be 256-byte aligned");
    const std::vector<Pack> packs = sm ? make_packs(problems, n) : std::vector<Pack>();
    std::vector<int64_t> pack_of(n, 0), b0_of(n, 0);
    for (size_t k = 0; k < packs.size(); ++k)
        for (size_t m = 0; m < packs[k].members.size(); ++m) {
            pack_of[packs[k].members[m]] = (int64_t)k;
            b0_of[packs[k].members[m]] = packs[k].b0[m];
        }

    // Tile counts and job order.
    std::vector<FwdArgs> fe(n), fo(n);
    std::vector<BwdArgs> bs(n), be(n), bo(n);
    std::vector<CopyJob> cp(n), cin(n), cout0(n), cout1(n);
    std::vector<double> fcost(n), bcost(n);
    char *ws = (char *)workspace;
    for (int64_t i = 0; i < n; ++i) {
        const fl_mlem_problem &q = problems[i];
        const int64_t k = pack_of[i], HW = q.H * q.W, nC = q.B * HW;
        // dense: this problem's regions; slice-major: its slices of its pack's
        float *sens = sm ? (float *)(ws + L.sens[k]) + b0_of[i] * HW : (float *)(ws + L.sens[i]);
        float *alt = sm ? (float *)(ws + L.alt[k]) + b0_of[i] * HW : (float *)(ws + L.alt[i]);
        float *alt2 = sm ? (float *)(ws + L.alt2[k]) + b0_of[i] * HW : nullptr;
        float *ratio = sm ? (float *)(ws + L.ratio[k]) + b0_of[i] : (float *)(ws + L.ratio[i]);
        const int G = fwd_angles_of(q), BI = bwd_rows_of(q);

        FwdArgs f{};
        f.atten = q.atten;
        f.em = q.em_cs;
        f.theta = q.theta;
        f.I = q.I;
        f.out = ratio;
        f.A = (int)q.A;
        f.B = (int)q.B;
        f.R = 1;
        f.H = (int)q.H;
        f.W = (int)q.W;
        f.r0 = 0;
        f.nref = 1;
        f.mode = kFwdRatio;
        f.ntw = (int)((q.W + kFW - 1) / kFW);
        f.n_groups = (int)((q.A + G - 1) / G);
        f.group_angles = G;
        if (sm) {
            // em[a] * ratio into the pack's (A, W, sum B) ratio at this problem's slices
            f.out_sm = 1;
            f.out_pitch = (int)packs[k].Bp;
            f.fold_em = 1;
        }
        fe[i] = f;
        fe[i].C = sm ? alt : q.C;
        fo[i] = f;
        fo[i].C = sm ? alt2 : alt;
        // a forward tile walks the strip's H rows for up to G angles
        fcost[i] = (double)q.H * (double)std::min<int64_t>(G, q.A);

        if (!sm) {
            BwdArgs b{};
            b.atten = q.atten;
            b.em = q.em_cs;
            b.theta = q.theta;
            b.A = (int)q.A;
            b.B = (int)q.B;
            b.R = 1;
            b.H = (int)q.H;
            b.W = (int)q.W;
            b.r0 = 0;
            b.nref = 1;
            b.nti = (int)((q.H + BI - 1) / BI);
            b.ntj = (int)((q.W + kBJ - 1) / kBJ);
            b.beta = (float)beta;
            b.delta = (float)delta;
            b.tile_rows = BI;
            bs[i] = b;
            bs[i].ratio = nullptr;
            bs[i].out = sens;
            bs[i].mode = kBwdSens;
            be[i] = b;
            be[i].ratio = ratio;
            be[i].mode = kBwdUpdate;
            be[i].C = q.C;
            be[i].sens = sens;
            be[i].out = alt;
            bo[i] = be[i];
            bo[i].C = alt;
            bo[i].out = q.C;
            // a back-projector tile gathers every angle's samples over BI x kBJ cells
            bcost[i] = (double)q.A * (double)std::min<int64_t>(BI, q.H) * (double)std::min<int64_t>(kBJ, q.W);
        }
        cp[i] = CopyJob{alt, q.C, nC};
        if (sm) {
            cin[i] = CopyJob{q.C, alt, nC};
            cout0[i] = CopyJob{alt, q.C, nC};
            cout1[i] = CopyJob{alt2, q.C, nC};
        }
    }
    const std::vector<int64_t> ford = order_by(n, fcost), bord = order_by(n, bcost);

    auto *m = new fl_mlem_multi();
    m->n = (int)n;
    m->sm = sm;
    m->staging.assign(L.tables, 0);
    char *h = m->staging.data();
    std::vector<uint32_t> fb(n + 1), bb(n + 1), cb(n + 1);
    uint64_t ft = 0, bt = 0, ct = 0;
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = ford[k], j = bord[k];
        ((FwdArgs *)(h + L.fwd[0]))[k] = fe[i];
        ((FwdArgs *)(h + L.fwd[1]))[k] = fo[i];
        fb[k] = (uint32_t)ft;
        ft += fwd_job_tiles(fe[i]);
        ((BwdArgs *)(h + L.sens_jobs))[k] = bs[j];
        ((BwdArgs *)(h + L.bwd[0]))[k] = be[j];
        ((BwdArgs *)(h + L.bwd[1]))[k] = bo[j];
        bb[k] = (uint32_t)bt;
        bt += sm ? 0 : bwd_job_tiles(be[j]);
        ((CopyJob *)(h + L.copy))[k] = cp[k];
        if (sm) {
            ((CopyJob *)(h + L.pack_in))[k] = cin[k];
            ((CopyJob *)(h + L.unpack[0]))[k] = cout0[k];
            ((CopyJob *)(h + L.unpack[1]))[k] = cout1[k];
        }
        cb[k] = (uint32_t)ct;
        ct += copy_job_tiles(cp[k]);
    }
    if (ft >= UINT32_MAX || bt >= UINT32_MAX || ct >= UINT32_MAX) {
        delete m;
        set_error("batch: too many tiles for one launch");
        return FL_ERR_INVALID_ARG;
    }
    fb[n] = (uint32_t)ft;
    bb[n] = (uint32_t)bt;
    cb[n] = (uint32_t)ct;
    std::copy(fb.begin(), fb.end(), (uint32_t *)(h + L.fwd_begin));
    std::copy(bb.begin(), bb.end(), (uint32_t *)(h + L.bwd_begin));
    std::copy(cb.begin(), cb.end(), (uint32_t *)(h + L.copy_begin));
    m->fwd_tiles = (uint32_t)ft;
    m->bwd_tiles = (uint32_t)bt;
    m->copy_tiles = (uint32_t)ct;
    m->fwd[0] = (FwdArgs *)(ws + L.fwd[0]);
    m->fwd[1] = (FwdArgs *)(ws + L.fwd[1]);
    m->bwd[0] = (BwdArgs *)(ws + L.bwd[0]);
    m->bwd[1] = (BwdArgs *)(ws + L.bwd[1]);
    m->sens = (BwdArgs *)(ws + L.sens_jobs);
    m->copy = (CopyJob *)(ws + L.copy);
    m->fwd_begin = (uint32_t *)(ws + L.fwd_begin);
    m->bwd_begin = (uint32_t *)(ws + L.bwd_begin);
    m->copy_begin = (uint32_t *)(ws + L.copy_begin);
    if (sm) {
        m->pack_in = (CopyJob *)(ws + L.pack_in);
        m->unpack[0] = (CopyJob *)(ws + L.unpack[0]);
        m->unpack[1] = (CopyJob *)(ws + L.unpack[1]);
        float *ones = (float *)(h + L.ones);
        int64_t max_a = 1;
        for (const Pack &pk : packs) max_a = std::max(max_a, pk.A);
        std::fill(ones, ones + max_a, 1.0f);
        // the grouped back-projection: per pack the sensitivity, even and odd finalize
        const int np = (int)packs.size();
        std::vector<SmPack> sp(np);
        std::vector<SmFinSpec> fin(3 * (size_t)np);
        for (int k = 0; k < np; ++k) {
            const Pack &pk = packs[k];
            float *sens = (float *)(ws + L.sens[k]), *c0 = (float *)(ws + L.alt[k]), *c1 = (float *)(ws + L.alt2[k]);
            sp[k] = SmPack{(const float *)(ws + L.atten_sm[k]), (const float *)(ws + L.ratio[k]),
                           (const float *)(ws + L.ones), pk.theta, pk.A, pk.Bp, pk.H, pk.W, pk.spl,
                           ws + L.scratch[k]};
            fin[k] = SmFinSpec{sens, kBwdSens, nullptr, nullptr, (float)beta, (float)delta};
            fin[np + k] = SmFinSpec{c1, kBwdUpdate, c0, sens, (float)beta, (float)delta};
            fin[2 * np + k] = SmFinSpec{c0, kBwdUpdate, c1, sens, (float)beta, (float)delta};
        }
        if (int e = sm_group_build(sp.data(), np, fin.data(), 3, h + L.sm_tables, &m->smg)) {
            delete m;
            return e;
        }
        m->smg.dev = ws + L.sm_tables;
    }

    hipStream_t s = (hipStream_t)stream;
    int e = FL_OK;
    auto fail = [&](int code) {
        delete m;
        return code;
    };
    if (hipGetDevice(&m->device) != hipSuccess) {
        set_error("hipGetDevice failed");
        return fail(FL_ERR_HIP);
    }
    if ((e = fwd_group_grid(&m->fwd_grid)) || (e = bwd_group_grid(&m->bwd_grid)) ||
        (e = copy_group_grid(&m->copy_grid)))
        return fail(e);
    const hipError_t ce = hipMemcpyAsync(ws, h, L.tables, hipMemcpyHostToDevice, s);
    if (ce != hipSuccess) {
        set_error(std::string("job table upload: ") + hipGetErrorString(ce));
        return fail(FL_ERR_HIP);
    }
    if (sm) {
        // the slice-major copies, each problem's em as its ratio, then
        // sens = max(B^T em, 1e-6) per slice   (atten_cuda.cpp:308-314)
        for (size_t k = 0; k < packs.size(); ++k) {
            const Pack &pk = packs[k];
            if (pk.Bp > pk.B) {
                // padding slices: zero attenuation, ratio and estimates, so that they stay 0
                const size_t at = L.atten_sm[k], end = L.alt2[k] + (size_t)(pk.Bp * pk.H * pk.W) * sizeof(float);
                const hipError_t me = hipMemsetAsync(ws + at, 0, end - at, s);
                if (me != hipSuccess) {
                    set_error(std::string("padding clear: ") + hipGetErrorString(me));
                    return fail(FL_ERR_HIP);
                }
            }
            for (size_t j = 0; j < pk.members.size(); ++j) {
                const fl_mlem_problem &q = problems[pk.members[j]];
                if ((e = launch_sm_transpose_pitched(q.atten, (float *)(ws + L.atten_sm[k]) + pk.b0[j], q.A, q.B,
                                                     q.H * q.W, pk.Bp, s)))
                    return fail(e);
                if ((e = launch_sm_fill_em((float *)(ws + L.ratio[k]) + pk.b0[j], q.em_cs, q.A, q.W, q.B, pk.Bp, s)))
                    return fail(e);
            }
        }
        if ((e = launch_sm_group(m->smg, 0, s))) return fail(e);
    } else {
        // sens_i = max(B_i^T 1, 1e-6)   (atten_cuda.cpp:308-314)
        if ((e = launch_bwd_group(m->sens, m->bwd_begin, m->n, m->bwd_tiles, m->bwd_grid, s))) return fail(e);
    }
    *plan = m;
    return FL_OK;
}

int fl_mlem_multi_run(fl_mlem_multi *m, int64_t n_iter, int use_graph, void *stream)
{
    set_error("");
    FL_REQUIRE(m != nullptr, "null plan");
    FL_REQUIRE(n_iter >= 0, "n_iter must be >= 0");
    int dev = -1;
    FL_HIP_CHECK(hipGetDevice(&dev));
    FL_REQUIRE(dev == m->device, "batch: the plan belongs to another device");
    if (n_iter == 0) return FL_OK;
    hipStream_t s = (hipStream_t)stream;
    const int64_t pairs = n_iter / 2;
    if (m->sm) {
        // the estimates into the first packed copy
        if (int e = launch_copy_group(m->pack_in, m->copy_begin, m->n, m->copy_tiles, m->copy_grid, s)) return e;
    }
    if (use_graph && pairs > 0) {
        FL_REQUIRE(s != nullptr, "batch: graph mode needs a non-default stream");
        if (!m->exec) {
            FL_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int e = enqueue_iteration(m, 0, s);
            if (!e) e = enqueue_iteration(m, 1, s);
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(s, &g);
            if (e || ee != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (!e) set_error(std::string("graph capture: ") + hipGetErrorString(ee));
                return e ? e : FL_ERR_HIP;
            }
            const hipError_t ie = hipGraphInstantiate(&m->exec, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(g);
                m->exec = nullptr;
                set_error(std::string("graph instantiate: ") + hipGetErrorString(ie));
                return FL_ERR_HIP;
            }
            m->graph = g;
        }
        for (int64_t k = 0; k < pairs; ++k) FL_HIP_CHECK(hipGraphLaunch(m->exec, s));
    } else {
        for (int64_t k = 0; k < pairs; ++k) {
            if (int e = enqueue_iteration(m, 0, s)) return e;
            if (int e = enqueue_iteration(m, 1, s)) return e;
        }
    }
    if (n_iter & 1) {
        // one more iteration into the second copy
        if (int e = enqueue_iteration(m, 0, s)) return e;
    }
    // the result back into C: from the workspace copy it ended in
    const CopyJob *back = m->sm ? m->unpack[n_iter & 1] : ((n_iter & 1) ? m->copy : nullptr);
    if (back)
        if (int e = launch_copy_group(back, m->copy_begin, m->n, m->copy_tiles, m->copy_grid, s)) return e;
    return FL_OK;
}

int fl_mlem_multi_destroy(fl_mlem_multi *m)
{
    set_error("");
    if (!m) return FL_OK;
    int rc = FL_OK;
    if (m->exec && hipGraphExecDestroy(m->exec) != hipSuccess) rc = FL_ERR_HIP;
    if (m->graph && hipGraphDestroy(m->graph) != hipSuccess) rc = FL_ERR_HIP;
    delete m;
    if (rc) set_error("batch: graph release failed");
    return rc;
}

}  // extern "C"
