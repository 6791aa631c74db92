// Back-projection (exact transpose of the forward projector's sampling) on a
// slice-major copy of the attenuation, for CDNA4.
//
// In the reference's layout atten[a,b,h,w] a back-projector that owns an image
// tile reads short row segments of every detector row it needs, and each
// sample's position and bilinear weights are computed once per slice.  The
// slice-major copy atten_sm[a,h,w,b] puts the B slices of one detector point in
// contiguous memory.  A wavefront here covers 64 * SPL slices (lane = SPL
// consecutive slices, one 4*SPL-byte load per point): the position, the clamping
// and the four weights of a detector point are the same for all of them and are
// computed once, and each lane adds into its own slices' accumulators, so no two
// lanes ever touch the same word.
//
// The detector points are partitioned by where the forward projector's sample
// of each lands, and every point is taken by exactly one work item:
//  * Interior blocks (sm_interior_body): points whose sample lies inside the
//    clamp box [0, H-1.001] x [0, W-1.001] and whose base cell (floor of the
//    position) is in one 4 x 4 block of base cells.  One wavefront per block:
//    at each angle lane 4c + q tests candidate q of cell c (the 2 x 2 lattice
//    points of the bounding box of the cell's preimage, which hold every point
//    of the cell), and the wave's hit mask lists that angle's points.  The hits
//    stream through a ring of 8 load slots across angles (the next angle with
//    hits is tested ahead, so the ring does not drain between angles): each
//    slot holds one point's attenuation and ratio loads and its weights, read
//    from the hit's lane.  The block's 5 x 5 window of voxels (cells plus their
//    lower/right neighbours) is 25 * SPL registers per lane, and a tree of
//    wave-uniform branches on the hit's cell selects its four target registers,
//    named at compile time.  The 4 x 4 part of the window is the block's own
//    output; its last row and column (the halo) go to a small buffer.
//  * Edge strips (sm_edge_body): points whose sample the reference clamps onto
//    the image border (.cu:419-420), split by the border they clamp onto
//    (row < 0: top, row > H-1.001: bottom, else col < 0: left, col > W-1.001:
//    right), by an 8-wide strip of border base cells, by 8-voxel depth bins and
//    by ranges of angles, so that no item holds a long chain of work.  A strip
//    enumerates the lattice points of the bounding box of its region's preimage,
//    64 per round, its four wavefronts taking every fourth angle, and adds into
//    its 2 x 9 (or 9 x 2) window of border voxels in registers with the clamped
//    position's weights; the wavefronts' windows are summed in LDS in a fixed
//    order.  Edge items and interior blocks share one launch (edge items first),
//    so that the longer edge items overlap the blocks.
// sm_finalize adds, per voxel and in a fixed order, the owning block's sum, the
// halos of the blocks above and to the left, and every edge strip covering the
// voxel, transposes back to the reference's (B, H, W) layout through LDS and
// applies the output mode (write, sensitivity, MLEM/OSL update, accumulate).
// No atomics anywhere: results are reproducible bit for bit.
#include "projector.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fl {

constexpr int kSmT = 8;                  // edge strips: 8 border base cells, 8-voxel depth bins
constexpr int kSmWin = kSmT + 1;         // edge window length along the border
constexpr int kSmEdgeVox = 2 * kSmWin;   // voxels of an edge strip's window
constexpr int kSmTB = 4;                 // interior block: 4 x 4 base cells
constexpr int kSmWB = kSmTB + 1;         // interior window side
constexpr int kSmWBVox = kSmWB * kSmWB;  // 25
constexpr int kSmHaloB = 2 * kSmTB + 1;  // halo of a block: row of 5, column of 4
constexpr uint32_t kSmOff = 0x7FFFFFF0u; // an offset past every buffer (reads 0)

enum SmKind { kSmTop = 0, kSmBottom = 1, kSmLeft = 2, kSmRight = 3 };

struct SmGeom {
    int A, B, H, W;
    // edge strips
    int nsi, nsj;    // strips along the left/right borders (rows) and top/bottom (columns)
    int nd;          // depth bins
    int nch;         // angle ranges per strip
    float reach;     // bound on |position - centre| over the detector, plus a margin
    // 64 * spl slices per wavefront, in Gi groups
    int spl, S, Gi;
    int nbi, nbj;    // 4 x 4 blocks over base cells 0..H-2 x 0..W-2
};

struct SmArgs {
    SmGeom g;
    const float *atten;  // (A, H, W, B)
    const float *ratio;  // (A, W, B), or null for ratio = 1
    const float *em;     // (A): one reference per line
    const float *theta;
    float *own;          // (Gi, nbi*4, nbj*4, S)
    float *halo;         // (Gi, nbi, nbj, kSmHaloB, S)
    float *edge;         // (Gi, 4, max(nsi, nsj), nd, nch, kSmEdgeVox, S)
    float *edge_sum;     // (Gi, 4, max(nsi, nsj), kSmEdgeVox, S)
};

__host__ __device__ inline int sm_edge_items(const SmGeom &g) { return (2 * g.nsj + 2 * g.nsi) * g.nd * g.nch; }

__host__ __device__ inline int64_t sm_edge_index(const SmGeom &g, int grp, int side, int t, int d, int ch)
{
    const int ns = g.nsi > g.nsj ? g.nsi : g.nsj;
    return ((((int64_t)grp * 4 + side) * ns + t) * g.nd + d) * g.nch + ch;
}

// Region of an edge strip in sample coordinates (row, col) and the origin of its
// voxel window.
struct SmItem {
    int side, t, d, ch;
    float r0, r1, c0, c1;
    int wy, wx;
};

__device__ __forceinline__ SmItem sm_item(const SmGeom &g, int k, float xmax, float ymax)
{
    // k runs over depth bins first (the bins next to the border hold most hits,
    // so their items start first), then strips (top, bottom, left, right), then
    // angle ranges
    SmItem it;
    const float cx = (g.W - 1) * 0.5f, cy = (g.H - 1) * 0.5f;
    const float rlo = cy - g.reach, rhi = cy + g.reach, clo = cx - g.reach, chi = cx + g.reach;
    const int nstrips = 2 * g.nsj + 2 * g.nsi;
    it.ch = k % g.nch;
    k /= g.nch;
    int st = k % nstrips;
    it.d = k / nstrips;
    const bool last_d = it.d == g.nd - 1;
    if (st < 2 * g.nsj) {
        const bool top = st < g.nsj;
        it.side = top ? kSmTop : kSmBottom;
        it.t = top ? st : st - g.nsj;
        if (top) {
            it.r0 = last_d ? rlo : -(float)(8 * (it.d + 1));
            it.r1 = -(float)(8 * it.d);
            it.wy = 0;
        } else {
            it.r0 = xmax + (float)(8 * it.d);
            it.r1 = last_d ? rhi : xmax + (float)(8 * (it.d + 1));
            it.wy = g.H - 2;
        }
        it.c0 = it.t == 0 ? clo : (float)(8 * it.t);
        it.c1 = it.t == g.nsj - 1 ? chi : (float)(8 * it.t + 8);
        it.wx = 8 * it.t;
        return it;
    }
    st -= 2 * g.nsj;
    const bool left = st < g.nsi;
    it.side = left ? kSmLeft : kSmRight;
    it.t = left ? st : st - g.nsi;
    it.r0 = (float)(8 * it.t);
    it.r1 = (float)(8 * it.t + 8);
    if (left) {
        it.c0 = last_d ? clo : -(float)(8 * (it.d + 1));
        it.c1 = -(float)(8 * it.d);
        it.wx = 0;
    } else {
        it.c0 = ymax + (float)(8 * it.d);
        it.c1 = last_d ? chi : ymax + (float)(8 * (it.d + 1));
        it.wx = g.W - 2;
    }
    it.wy = 8 * it.t;
    return it;
}

// Depth bin of a distance v >= 0 beyond the clamp box.
__device__ __forceinline__ int sm_depth(float v, int nd) { return min((int)(v * 0.125f), nd - 1); }

template <int N>
struct alignas(4 * N) SmVec {
    float v[N];
};

template <int N, int AUX>
__device__ __forceinline__ SmVec<N> sm_load(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    SmVec<N> o;
    if constexpr (N == 1) {
        o.v[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX));
    } else if constexpr (N == 2) {
        const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX);
        o.v[0] = __uint_as_float(x[0]);
        o.v[1] = __uint_as_float(x[1]);
    } else {
        const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
        o.v[0] = __uint_as_float(x[0]);
        o.v[1] = __uint_as_float(x[1]);
        o.v[2] = __uint_as_float(x[2]);
        o.v[3] = __uint_as_float(x[3]);
    }
    return o;
}

// ---------------------------------------------------------------------------
// Edge strips.  In strip coordinates a voxel of the window is (across, along):
// across 0/1 is the window's first/second row (top and bottom strips) or
// column (left and right strips), along 0..8 runs with the border.  A hit whose
// base voxel is at along = P adds to (0, P), (0, P + 1), (1, P), (1, P + 1); the
// window is 18 * SPL registers per lane, and each hit is dispatched to the
// registers of its P through a tree of wave-uniform branches.  The kSmEW
// wavefronts of a workgroup share one item and take its angles in turn; their
// windows are summed in wavefront order through LDS at the end.
// ---------------------------------------------------------------------------
constexpr int kSmEW = 4;  // wavefronts per edge item

template <int P, int SPL>
__device__ __forceinline__ void sm_edge_acc(SmVec<SPL> *acc, const SmVec<SPL> &v, float a00, float a01, float a10,
                                            float a11)
{
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
        acc[P].v[j] = fmaf(a00, v.v[j], acc[P].v[j]);
        acc[P + 1].v[j] = fmaf(a01, v.v[j], acc[P + 1].v[j]);
        acc[kSmWin + P].v[j] = fmaf(a10, v.v[j], acc[kSmWin + P].v[j]);
        acc[kSmWin + P + 1].v[j] = fmaf(a11, v.v[j], acc[kSmWin + P + 1].v[j]);
    }
    // a marker that differs per position (see sm_acc_cell)
    asm volatile("; strip position %0" ::"n"(P));
}

template <int LO, int HI, int SPL>
__device__ __forceinline__ void sm_edge_dispatch(int pos, SmVec<SPL> *acc, const SmVec<SPL> &v, float a00, float a01,
                                                 float a10, float a11)
{
    if constexpr (LO == HI) {
        sm_edge_acc<LO, SPL>(acc, v, a00, a01, a10, a11);
    } else {
        constexpr int MID = (LO + HI) / 2;
        if (pos <= MID)
            sm_edge_dispatch<LO, MID, SPL>(pos, acc, v, a00, a01, a10, a11);
        else
            sm_edge_dispatch<MID + 1, HI, SPL>(pos, acc, v, a00, a01, a10, a11);
    }
}

// Per-lane fields of one round of candidates: offsets, position along the strip
// and the four weights in strip coordinates (the emission factor folded in).
struct SmEdgeCand {
    uint32_t ao, ro;
    int pos;
    float w00, w01, w10, w11;
};

// Workgroup `id` of one problem's edge launch.
template <int SPL, bool RATIO>
__device__ __forceinline__ void sm_edge_body(const SmArgs &p, uint32_t id)
{
    constexpr int S = 64 * SPL;
    constexpr int RING = 8;
    __shared__ SmVec<SPL> win[kSmEdgeVox * 64];
    const SmGeom &g = p.g;
    // items in dispatch order (the dispatcher deals consecutive workgroups to
    // the XCDs in turn), slice groups innermost
    const int grp = (int)(id % (uint32_t)g.Gi);
    const int k = (int)(id / (uint32_t)g.Gi);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b0 = grp * S + lane * SPL;  // this lane's first slice (B is a multiple of SPL)
    const float xmax = g.H - 1.001f, ymax = g.W - 1.001f;
    const float cx = (g.W - 1) * 0.5f, cy = (g.H - 1) * 0.5f;
    const SmItem it = sm_item(g, k, xmax, ymax);
    const int64_t HWB = (int64_t)g.H * g.W * g.B;
    const int64_t WB = (int64_t)g.W * g.B;
    // lanes past the last slice read through an offset past every buffer (0)
    const uint32_t voff = b0 < g.B ? (uint32_t)b0 * 4u : kSmOff;
    const bool rows = it.side == kSmTop || it.side == kSmBottom;
    const int a_begin = (int)((int64_t)g.A * it.ch / g.nch), a_end = (int)((int64_t)g.A * (it.ch + 1) / g.nch);

    SmVec<SPL> acc[kSmEdgeVox];
#pragma unroll
    for (int v = 0; v < kSmEdgeVox; ++v)
#pragma unroll
        for (int j = 0; j < SPL; ++j) acc[v].v[j] = 0.0f;

    // Every sample lies in the disc of radius `reach` about the centre (the
    // rotation preserves distances): clip the region's columns to the disc's
    // chord over its rows, and its rows to the chord over its columns.
    float r0 = fminf(it.r0, it.r1), r1 = fmaxf(it.r0, it.r1), c0 = fminf(it.c0, it.c1), c1 = fmaxf(it.c0, it.c1);
    {
        const float R2 = g.reach * g.reach;
        const float dr = fmaxf(0.0f, fmaxf(r0 - cy, cy - r1)), dc = fmaxf(0.0f, fmaxf(c0 - cx, cx - c1));
        const float hc = sqrtf(fmaxf(R2 - dr * dr, 0.0f)), hr = sqrtf(fmaxf(R2 - dc * dc, 0.0f));
        c0 = fmaxf(c0, cx - hc);
        c1 = fminf(c1, cx + hc);
        r0 = fmaxf(r0, cy - hr);
        r1 = fminf(r1, cy + hr);
    }
    const bool reachable = r0 <= r1 && c0 <= c1;

    // ---- the test cursor: this wavefront's angles a_begin + wave + n * kSmEW,
    // each enumerating the lattice points of the bounding box of the preimage of
    // the item's region (inverse rotation of its corners, one lattice step of
    // margin, clipped to the detector), 64 per round
    int ta = reachable ? a_begin + wave : a_end;  // angle under test
    int tb = 0, ncand = 0;                        // next candidate of that angle, candidates
    int h0 = 0, w0 = 0, nw = 1;
    float tc = 1.0f, ts = 0.0f, tem = 0.0f, inv_nw = 1.0f;
    auto open_angle = [&]() {  // bounding box of angle ta; ncand = 0 when empty
        const float th = finite_or_zero(scalar_read(p.theta, ta));
        tc = cosf(-th);
        ts = sinf(-th);
        tem = scalar_read(p.em, ta);
        float hlo = 3.0e38f, hhi = -3.0e38f, wlo = 3.0e38f, whi = -3.0e38f;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
            const float rr = ((qq & 1) ? r1 : r0) - cy;
            const float cc = ((qq & 2) ? c1 : c0) - cx;
            const float hh = tc * rr + ts * cc, ww = -ts * rr + tc * cc;
            hlo = fminf(hlo, hh);
            hhi = fmaxf(hhi, hh);
            wlo = fminf(wlo, ww);
            whi = fmaxf(whi, ww);
        }
        h0 = max(0, (int)floorf(hlo + cy) - 1);
        const int h1 = min(g.H - 1, (int)ceilf(hhi + cy) + 1);
        w0 = max(0, (int)floorf(wlo + cx) - 1);
        const int w1 = min(g.W - 1, (int)ceilf(whi + cx) + 1);
        nw = max(w1 - w0 + 1, 1);
        ncand = (h0 > h1 || w0 > w1) ? 0 : (h1 - h0 + 1) * nw;
        inv_nw = 1.0f / (float)nw;
        tb = 0;
    };
    // one round of candidates: lane = one lattice point; returns the hit mask
    auto test = [&](SmEdgeCand &m) -> uint64_t {
        const int kc = tb + lane;
        int qh = (int)(((float)kc + 0.5f) * inv_nw);
        int qw = kc - qh * nw;
        if (qw < 0) {
            --qh;
            qw += nw;
        } else if (qw >= nw) {
            ++qh;
            qw -= nw;
        }
        const int h = h0 + qh, w = w0 + qw;
        bool hit = kc < ncand;
        // the forward projector's position arithmetic, unclamped
        const float hc = (float)h - cy, wc = (float)w - cx;
        const float xu = fmaf(tc, hc, fmaf(-ts, wc, cy));  // row
        const float yu = fmaf(ts, hc, fmaf(tc, wc, cx));   // col
        const bool rin = xu >= 0.0f && xu <= xmax;
        const float xc = clampf(xu, 0.0f, xmax), yc = clampf(yu, 0.0f, ymax);
        const int y0 = (int)xc, x0 = (int)yc;
        const float fy = __builtin_amdgcn_fractf(xc), fx = __builtin_amdgcn_fractf(yc);
        if (it.side == kSmTop) {
            hit = hit && xu < 0.0f && (x0 >> 3) == it.t && sm_depth(-xu, g.nd) == it.d;
        } else if (it.side == kSmBottom) {
            hit = hit && xu > xmax && (x0 >> 3) == it.t && sm_depth(xu - xmax, g.nd) == it.d;
        } else if (it.side == kSmLeft) {
            hit = hit && rin && yu < 0.0f && (y0 >> 3) == it.t && sm_depth(-yu, g.nd) == it.d;
        } else {
            hit = hit && rin && yu > ymax && (y0 >> 3) == it.t && sm_depth(yu - ymax, g.nd) == it.d;
        }
        m.ao = (uint32_t)((h * g.W + w) * g.B) * 4u;
        m.ro = (uint32_t)(w * g.B) * 4u;
        m.pos = rows ? x0 - it.wx : y0 - it.wy;
        const float e1 = tem * (1.0f - fy), e0 = tem * fy;
        // image weights: (y0, x0), (y0, x0 + 1), (y0 + 1, x0), (y0 + 1, x0 + 1);
        // a column strip runs along rows, so its (0, P + 1) is (y0 + 1, x0)
        const float i01 = e1 * fx, i10 = e0 * (1.0f - fx);
        m.w00 = e1 * (1.0f - fx);
        m.w01 = rows ? i01 : i10;
        m.w10 = rows ? i10 : i01;
        m.w11 = e0 * fx;
        tb += 64;
        return __ballot(hit);
    };

    // ---- hits streamed through a ring of RING load slots ---------------------
    SmEdgeCand cur{}, nxt{};
    uint64_t mc = 0, mx = 0;
    int ac = 0, ax = 0;
    __amdgpu_buffer_rsrc_t ra = make_rsrc(p.atten, 0), rq = ra;
    auto find_next = [&]() {
        while (mx == 0 && ta < a_end) {
            if (tb == 0) open_angle();
            if (tb < ncand) {
                ax = ta;
                mx = test(nxt);
            }
            if (tb >= ncand) {
                ta += kSmEW;
                tb = 0;
            }
        }
    };

    SmVec<SPL> va[RING], vr[RING];
    int sp[RING];  // position of the hit in each slot, -1: empty
    float s00[RING], s01[RING], s10[RING], s11[RING];
    auto issue = [&](int r) {
        if (mc == 0 && mx != 0) {
            cur = nxt;
            mc = mx;
            ac = ax;
            mx = 0;
            ra = make_rsrc(p.atten + (int64_t)ac * HWB, (int)HWB);
            if constexpr (RATIO) rq = make_rsrc(p.ratio + (int64_t)ac * WB, (int)WB);
        }
        uint32_t ao = kSmOff, ro = kSmOff;
        sp[r] = -1;
        if (mc != 0) {
            const int l = __builtin_ctzll(mc);
            mc &= mc - 1;
            sp[r] = __builtin_amdgcn_readlane(cur.pos, l);
            s00[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.w00), l));
            s01[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.w01), l));
            s10[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.w10), l));
            s11[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.w11), l));
            ao = (uint32_t)__builtin_amdgcn_readlane((int)cur.ao, l);
            ro = (uint32_t)__builtin_amdgcn_readlane((int)cur.ro, l);
        }
        va[r] = sm_load<SPL, 2>(ra, voff + ao);
        if constexpr (RATIO) vr[r] = sm_load<SPL, 0>(rq, voff + ro);
    };

    find_next();
#pragma unroll
    for (int r = 0; r < RING; ++r) issue(r);
    for (;;) {
        find_next();
        bool busy = false;
#pragma unroll
        for (int r = 0; r < RING; ++r) {
            if (sp[r] >= 0) {  // wave-uniform
                busy = true;
                SmVec<SPL> v = va[r];
                if constexpr (RATIO) {
#pragma unroll
                    for (int j = 0; j < SPL; ++j) v.v[j] = v.v[j] * vr[r].v[j];
                }
                sm_edge_dispatch<0, kSmT - 1, SPL>(sp[r], acc, v, s00[r], s01[r], s10[r], s11[r]);
            }
            issue(r);
        }
        if (!busy) break;
    }

    // ---- the wavefronts' windows summed in wavefront order; entry q = across * 9 + along
    for (int u = 0; u < kSmEW; ++u) {
        if (wave == u) {
#pragma unroll
            for (int v = 0; v < kSmEdgeVox; ++v) {
                SmVec<SPL> x = acc[v];
                if (u > 0) {
                    const SmVec<SPL> y = win[v * 64 + lane];
#pragma unroll
                    for (int j = 0; j < SPL; ++j) x.v[j] = y.v[j] + x.v[j];
                }
                win[v * 64 + lane] = x;
            }
        }
        __syncthreads();
    }
    SmVec<SPL> *e = reinterpret_cast<SmVec<SPL> *>(p.edge + sm_edge_index(g, grp, it.side, it.t, it.d, it.ch) *
                                                                kSmEdgeVox * S);
    for (int q = wave; q < kSmEdgeVox; q += kSmEW) e[q * 64 + lane] = win[q * 64 + lane];
}


// Sum of a strip's windows over its depth bins and angle ranges, in that fixed
// order: one workgroup per (slice group, side, strip, window entry).
template <int SPL>
__device__ __forceinline__ void sm_edge_reduce_body(const SmArgs &p, uint32_t wg)
{
    const SmGeom &g = p.g;
    const int ns = g.nsi > g.nsj ? g.nsi : g.nsj;
    int id = (int)wg;
    const int q = id % kSmEdgeVox;
    id /= kSmEdgeVox;
    const int t = id % ns;
    id /= ns;
    const int side = id % 4, grp = id / 4;
    if (t >= (side < 2 ? g.nsj : g.nsi)) return;
    const int lane = threadIdx.x;
    const int n = g.nd * g.nch;
    const SmVec<SPL> *src = reinterpret_cast<const SmVec<SPL> *>(p.edge) +
                            (sm_edge_index(g, grp, side, t, 0, 0) * kSmEdgeVox + q) * 64 + lane;
    SmVec<SPL> acc = src[0];
    for (int i = 1; i < n; ++i) {
        const SmVec<SPL> x = src[(int64_t)i * kSmEdgeVox * 64];
#pragma unroll
        for (int j = 0; j < SPL; ++j) acc.v[j] += x.v[j];
    }
    reinterpret_cast<SmVec<SPL> *>(p.edge_sum)[((((int64_t)grp * 4 + side) * ns + t) * kSmEdgeVox + q) * 64 + lane] =
        acc;
}

template <int SPL>
__global__ __launch_bounds__(64) void sm_edge_reduce_kernel(SmArgs p)
{
    sm_edge_reduce_body<SPL>(p, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Interior blocks.
// ---------------------------------------------------------------------------
// acc[v] += w * P for the four window voxels of base cell C (row C / 4, column
// C % 4 of the block): register indices fixed at compile time.
template <int C, int SPL>
__device__ __forceinline__ void sm_acc_cell(SmVec<SPL> *acc, const SmVec<SPL> &P, float a00, float a01, float a10,
                                            float a11)
{
    constexpr int v = (C >> 2) * kSmWB + (C & 3);
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
        acc[v].v[j] = fmaf(a00, P.v[j], acc[v].v[j]);
        acc[v + 1].v[j] = fmaf(a01, P.v[j], acc[v + 1].v[j]);
        acc[v + kSmWB].v[j] = fmaf(a10, P.v[j], acc[v + kSmWB].v[j]);
        acc[v + kSmWB + 1].v[j] = fmaf(a11, P.v[j], acc[v + kSmWB + 1].v[j]);
    }
    // A marker that differs per cell: without it the compiler merges the 16
    // cells' identical code into one block indexing acc with a variable, which
    // moves acc out of registers into scratch memory.
    asm volatile("; window cell %0" ::"n"(C));
}

// Binary tree of wave-uniform branches on the cell index (16 cells: 4 levels).
template <int LO, int HI, int SPL>
__device__ __forceinline__ void sm_acc_dispatch(int c, SmVec<SPL> *acc, const SmVec<SPL> &P, float a00, float a01,
                                                float a10, float a11)
{
    if constexpr (LO == HI) {
        sm_acc_cell<LO, SPL>(acc, P, a00, a01, a10, a11);
    } else {
        constexpr int MID = (LO + HI) / 2;
        if (c <= MID)
            sm_acc_dispatch<LO, MID, SPL>(c, acc, P, a00, a01, a10, a11);
        else
            sm_acc_dispatch<MID + 1, HI, SPL>(c, acc, P, a00, a01, a10, a11);
    }
}

// Per-lane fields of one tested angle: the hit's atten and ratio byte offsets
// and its four weights (the emission factor folded in).
struct SmCand {
    uint32_t ao, ro;
    float w00, w01, w10, w11;
};

// Workgroup `id` of one problem's interior launch.
template <int SPL, bool RATIO>
__device__ __forceinline__ void sm_interior_body(const SmArgs &p, uint32_t id)
{
    constexpr int S = 64 * SPL;
    // hits whose loads are in flight
    constexpr int RING = 8;
    const SmGeom &g = p.g;
    const int nint = g.nbi * g.nbj;
    const int grp = (int)(id / (uint32_t)nint);
    const int k = (int)(id % (uint32_t)nint);
    const int bi = k / g.nbj, bj = k % g.nbj;
    const int lane = threadIdx.x & 63;
    const int b0 = grp * S + lane * SPL;  // this lane's first slice (B is a multiple of SPL)
    const float xmax = g.H - 1.001f, ymax = g.W - 1.001f;
    const float cx = (g.W - 1) * 0.5f, cy = (g.H - 1) * 0.5f;
    const int64_t HWB = (int64_t)g.H * g.W * g.B;
    const int64_t WB = (int64_t)g.W * g.B;
    const uint32_t voff = b0 < g.B ? (uint32_t)b0 * 4u : kSmOff;
    // lane 4c + q: candidate q of cell c = (c / 4, c % 4) of the block
    const int cell = lane >> 2, q = lane & 3;
    const int Y = bi * kSmTB + (cell >> 2), X = bj * kSmTB + (cell & 3);
    const float rr = (float)Y + 0.5f - cy, cc = (float)X + 0.5f - cx;

    SmVec<SPL> acc[kSmWBVox];
#pragma unroll
    for (int v = 0; v < kSmWBVox; ++v)
#pragma unroll
        for (int j = 0; j < SPL; ++j) acc[v].v[j] = 0.0f;

    // Candidate q of this lane's cell at angle a: the lattice points of the
    // bounding box of the cell's preimage (a rotated unit square), widened for
    // rounding, are the 2 x 2 points from (ha, wa).  Returns the wave's hit mask.
    auto test = [&](int a, SmCand &m) -> uint64_t {
        const float th = finite_or_zero(scalar_read(p.theta, a));
        const float c = cosf(-th), s = sinf(-th);
        const float em = scalar_read(p.em, a);
        const float hs = fmaf(c, rr, s * cc) + cy, ws = fmaf(-s, rr, c * cc) + cx;
        const float rad = 0.5f * (fabsf(c) + fabsf(s)) + 1e-3f;
        const int ha = (int)ceilf(hs - rad), wa = (int)ceilf(ws - rad);
        const int h = ha + (q >> 1), w = wa + (q & 1);
        const float hc = (float)h - cy, wc = (float)w - cx;
        const float xu = fmaf(c, hc, fmaf(-s, wc, cy));  // the forward's arithmetic
        const float yu = fmaf(s, hc, fmaf(c, wc, cx));
        const bool hit = h >= 0 && h < g.H && w >= 0 && w < g.W && xu >= 0.0f && xu <= xmax && yu >= 0.0f &&
                         yu <= ymax && (int)xu == Y && (int)yu == X;
        const float fy = __builtin_amdgcn_fractf(xu), fx = __builtin_amdgcn_fractf(yu);
        const float e1 = em * (1.0f - fy), e0 = em * fy;
        m.ao = (uint32_t)((h * g.W + w) * g.B) * 4u;
        m.ro = (uint32_t)(w * g.B) * 4u;
        m.w00 = e1 * (1.0f - fx);
        m.w01 = e1 * fx;
        m.w10 = e0 * (1.0f - fx);
        m.w11 = e0 * fx;
        return __ballot(hit);
    };

    // The hits are streamed through a ring of RING load slots across angles:
    // `cur` is the angle whose hits are being issued, `nxt` the next angle with
    // hits, tested ahead so that the ring does not drain between angles.
    SmCand cur{}, nxt{};
    uint64_t mc = 0, mx = 0;
    int ac = 0, ax = 0, an = 0;
    __amdgpu_buffer_rsrc_t ra = make_rsrc(p.atten, 0), rq = ra;
    auto find_next = [&]() {
        while (mx == 0 && an < g.A) {
            ax = an;
            mx = test(an, nxt);
            ++an;
        }
    };

    SmVec<SPL> va[RING], vr[RING];
    int sc[RING];  // cell of the hit in each slot, -1: empty
    float s00[RING], s01[RING], s10[RING], s11[RING];

    // the next hit into slot r (loads through an offset past the buffers when none is left)
    auto issue = [&](int r) {
        if (mc == 0 && mx != 0) {
            cur = nxt;
            mc = mx;
            ac = ax;
            mx = 0;
            ra = make_rsrc(p.atten + (int64_t)ac * HWB, (int)HWB);
            if constexpr (RATIO) rq = make_rsrc(p.ratio + (int64_t)ac * WB, (int)WB);
        }
        uint32_t ao = kSmOff, ro = kSmOff;
        sc[r] = -1;
        if (mc != 0) {
            const int l = __builtin_ctzll(mc);
            mc &= mc - 1;
            sc[r] = l >> 2;
            s00[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.w00), l));
            s01[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.w01), l));
            s10[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.w10), l));
            s11[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.w11), l));
            // the point's offset in this lane's vector address (a VALU-written
            // SGPR read by a memory instruction would cost 5 wait states)
            ao = (uint32_t)__builtin_amdgcn_readlane((int)cur.ao, l);
            ro = (uint32_t)__builtin_amdgcn_readlane((int)cur.ro, l);
        }
        va[r] = sm_load<SPL, 2>(ra, voff + ao);
        if constexpr (RATIO) vr[r] = sm_load<SPL, 0>(rq, voff + ro);
    };

    find_next();
#pragma unroll
    for (int r = 0; r < RING; ++r) issue(r);
    for (;;) {
        find_next();
        bool busy = false;
#pragma unroll
        for (int r = 0; r < RING; ++r) {
            if (sc[r] >= 0) {  // wave-uniform
                busy = true;
                SmVec<SPL> P = va[r];
                if constexpr (RATIO) {
#pragma unroll
                    for (int j = 0; j < SPL; ++j) P.v[j] = P.v[j] * vr[r].v[j];
                }
                sm_acc_dispatch<0, kSmTB * kSmTB - 1, SPL>(sc[r], acc, P, s00[r], s01[r], s10[r], s11[r]);
            }
            issue(r);
        }
        if (!busy) break;
    }

    // ---- the window: own 4 x 4 voxels, then the halo row and column ---------
    const int HO = g.nbi * kSmTB, WO = g.nbj * kSmTB;
    SmVec<SPL> *own = reinterpret_cast<SmVec<SPL> *>(p.own + (int64_t)grp * HO * WO * S);
#pragma unroll
    for (int yy = 0; yy < kSmTB; ++yy)
#pragma unroll
        for (int xx = 0; xx < kSmTB; ++xx)
            own[((int64_t)(bi * kSmTB + yy) * WO + (bj * kSmTB + xx)) * 64 + lane] = acc[yy * kSmWB + xx];
    SmVec<SPL> *hal = reinterpret_cast<SmVec<SPL> *>(p.halo + (((int64_t)grp * g.nbi + bi) * g.nbj + bj) * kSmHaloB * S);
#pragma unroll
    for (int xx = 0; xx < kSmWB; ++xx) hal[xx * 64 + lane] = acc[kSmTB * kSmWB + xx];
#pragma unroll
    for (int yy = 0; yy < kSmTB; ++yy) hal[(kSmWB + yy) * 64 + lane] = acc[yy * kSmWB + kSmTB];
}

// Interior blocks and edge strips in one launch: workgroups [0, nq8) take four
// interior blocks each, one per wavefront, in the XCD-contiguous order of the
// interior's own launch (nq8: the quads rounded up to a multiple of the XCD
// count), the rest the edge items in their dispatch order.  The short edge
// items then fill the slots that the interior's last blocks leave idle,
// instead of starting only after all of them.
template <int SPL, bool RATIO>
__device__ __forceinline__ void sm_fused_body(const SmArgs &p, uint32_t wg, uint32_t nwg)
{
    const SmGeom &g = p.g;
    const uint32_t nint = (uint32_t)g.Gi * g.nbi * g.nbj;
    const uint32_t nq8 = ((nint + kSmEW - 1) / kSmEW + kNumXcd - 1) / kNumXcd * kNumXcd;
    if (wg < nq8) {
        const uint32_t blk = xcd_remap(wg, nq8) * kSmEW + (threadIdx.x >> 6);
        if (blk < nint) sm_interior_body<SPL, RATIO>(p, blk);
        return;
    }
    const uint32_t item = wg - nq8;
    if (item < (uint32_t)g.Gi * (uint32_t)sm_edge_items(g)) sm_edge_body<SPL, RATIO>(p, item);
}

template <int SPL, bool RATIO>
__global__ __launch_bounds__(64 * kSmEW) void sm_fused_kernel(SmArgs p)
{
    sm_fused_body<SPL, RATIO>(p, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Finalize: per voxel the fixed-order sum of every item's contribution, back to
// (B, H, W), with the output mode applied.  One workgroup: one image row y,
// 64 columns, 64 slices; first lane = slice, then lane = column.
// ---------------------------------------------------------------------------
struct SmFin {
    SmGeom g;
    const float *own, *halo, *edge_sum;
    float *out;          // (B, H, W)
    int mode;            // BwdMode
    const float *C;      // update: (B, H, W)
    const float *sens;   // update
    float beta, delta;
};

__device__ __forceinline__ float sm_voxel_sum(const SmFin &f, int b, int y, int x)
{
    const SmGeom &g = f.g;
    const int HO = g.nbi * kSmTB, WO = g.nbj * kSmTB;
    const int gi = b / g.S, li = b % g.S;
    const int bi = y / kSmTB, bj = x / kSmTB, ry = y % kSmTB, rx = x % kSmTB;
    float v = 0.0f;
    if (bi < g.nbi && bj < g.nbj) v = f.own[(((int64_t)gi * HO + y) * WO + x) * g.S + li];
    const float *H0 = f.halo + (int64_t)gi * g.nbi * g.nbj * kSmHaloB * g.S;
    // bottom halo of the block above: its window row 4
    if (ry == 0 && bi >= 1 && bi - 1 < g.nbi) {
        if (bj < g.nbj) v += H0[(((int64_t)(bi - 1) * g.nbj + bj) * kSmHaloB + rx) * g.S + li];
        if (rx == 0 && bj >= 1 && bj - 1 < g.nbj)
            v += H0[(((int64_t)(bi - 1) * g.nbj + bj - 1) * kSmHaloB + kSmTB) * g.S + li];
    }
    // right halo of the block to the left: its window column 4
    if (rx == 0 && bj >= 1 && bj - 1 < g.nbj && bi < g.nbi)
        v += H0[(((int64_t)bi * g.nbj + bj - 1) * kSmHaloB + kSmWB + ry) * g.S + li];
    // edge strips covering the voxel, side by side
    const int ns = g.nsi > g.nsj ? g.nsi : g.nsj;
    for (int side = 0; side < 4; ++side) {
        const bool rows = side < 2;
        const int along = rows ? x : y;              // position along the border
        const int across = rows ? (side == kSmTop ? y : y - (g.H - 2)) : (side == kSmLeft ? x : x - (g.W - 2));
        if (across < 0 || across > 1) continue;
        const int nstrip = rows ? g.nsj : g.nsi;
        for (int dt = 0; dt < 2; ++dt) {
            const int t = (along >> 3) - dt;
            const int pos = along - 8 * t;           // 0..8 within the strip's window
            if (t < 0 || t >= nstrip || pos > kSmT) continue;
            const int qq = across * kSmWin + pos;
            v += f.edge_sum[((((int64_t)gi * 4 + side) * ns + t) * kSmEdgeVox + qq) * g.S + li];
        }
    }
    return v;
}

// Workgroup (columns 64*bx.., row y, slices 64*c64..) of one problem's finalize.
__device__ __forceinline__ void sm_finalize_body(const SmFin &f, int bx, int y, int c64)
{
    __shared__ float tile[64 * 65];
    const SmGeom &g = f.g;
    const int x0 = bx * 64;
    const int lane = threadIdx.x;
    // phase 1: lane = slice; voxels x0..x0+63 of row y
    const int nx = min(64, g.W - x0);
    const int bl = c64 * 64 + lane;
    if (bl < g.B)
        for (int i = 0; i < nx; ++i) tile[lane * 65 + i] = sm_voxel_sum(f, bl, y, x0 + i);
    __syncthreads();
    // phase 2: lane = column
    const int x = x0 + lane;
    if (x >= g.W) return;
    const int64_t HW = (int64_t)g.H * g.W;
    const int nb = min(64, g.B - c64 * 64);
    for (int sl = 0; sl < nb; ++sl) {
        const int bb = c64 * 64 + sl;
        const float back = tile[sl * 65 + lane];
        const int64_t o = (int64_t)bb * HW + (int64_t)y * g.W + x;
        if (f.mode == kBwdWrite) {
            f.out[o] = back;
        } else if (f.mode == kBwdAccumulate) {
            f.out[o] += back;
        } else if (f.mode == kBwdSens) {
            f.out[o] = back < 1e-6f ? 1e-6f : back;  // sens.clamp_min_(1e-6)
        } else {
            // C.mul_(back).div_(sens + beta*grad); C.clamp_min_(0)  (atten_cuda.cpp:334-341)
            const float Cv = f.C[o];
            float den = f.sens[o];
            if (f.beta > 0.0f) den = den + f.beta * huber_at(f.C + (int64_t)bb * HW, y, x, g.H, g.W, f.delta);
            const float v = (Cv * back) / den;
            f.out[o] = v < 0.0f ? 0.0f : v;
        }
    }
}

__global__ __launch_bounds__(64) void sm_finalize_kernel(SmFin f)
{
    sm_finalize_body(f, blockIdx.x, blockIdx.y, blockIdx.z);
}

// ---------------------------------------------------------------------------
// Grouped launches over many problems (batched MLEM): job j of a launch owns
// the workgroups [begin[j], begin[j+1]) and runs them exactly as its own
// launch would.  One launch per kernel and slices-per-lane class.
// ---------------------------------------------------------------------------
// Job j owns workgroups [begin[j], begin[j+1]).
template <int SPL>
__global__ __launch_bounds__(64 * kSmEW) void sm_fused_group_kernel(const SmArgs *jobs, const uint32_t *begin, int n)
{
    const int j = find_job(begin, n, blockIdx.x);
    const SmArgs p = scalar_record(jobs + j);
    const uint32_t b0 = scalar_read_u32(begin, j);
    sm_fused_body<SPL, true>(p, blockIdx.x - b0, scalar_read_u32(begin, j + 1) - b0);
}

template <int SPL>
__global__ __launch_bounds__(64) void sm_edge_reduce_group_kernel(const SmArgs *jobs, const uint32_t *begin, int n)
{
    const int j = find_job(begin, n, blockIdx.x);
    const SmArgs p = scalar_record(jobs + j);
    sm_edge_reduce_body<SPL>(p, blockIdx.x - scalar_read_u32(begin, j));
}

__global__ __launch_bounds__(64) void sm_finalize_group_kernel(const SmFin *jobs, const uint32_t *begin, int n)
{
    const int j = find_job(begin, n, blockIdx.x);
    const SmFin f = scalar_record(jobs + j);
    uint32_t id = blockIdx.x - scalar_read_u32(begin, j);
    const uint32_t nbx = (uint32_t)(f.g.W + 63) / 64;
    const int bx = (int)(id % nbx);
    id /= nbx;
    sm_finalize_body(f, bx, (int)(id % (uint32_t)f.g.H), (int)(id / (uint32_t)f.g.H));
}

// ---------------------------------------------------------------------------
// (N, B, L) -> (N, L, B): the slice-major copy.  64 x 64 tiles through LDS.
// ---------------------------------------------------------------------------
// dst rows hold Bp >= B slices (a pitched copy writes slices 0..B-1 of them).
__global__ __launch_bounds__(256) void sm_transpose_kernel(const float *src, float *dst, int B, int64_t L, int Bp)
{
    __shared__ float tile[64 * 65];
    const int64_t l0 = (int64_t)blockIdx.x * 64;
    const int b0 = blockIdx.y * 64;
    const int64_t n = blockIdx.z;
    const float *s = src + n * (int64_t)B * L;
    float *d = dst + n * (int64_t)Bp * L;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int bb = b0 + r;
        const int64_t l = l0 + tx;
        tile[r * 65 + tx] = (bb < B && l < L) ? s[(int64_t)bb * L + l] : 0.0f;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int64_t l = l0 + r;
        const int bb = b0 + tx;
        if (l < L && bb < B) d[l * Bp + bb] = tile[tx * 65 + r];
    }
}

// ratio[(a * W + w) * Bp + b] = em[a], b < B: one workgroup per (w, a).
__global__ __launch_bounds__(64) void sm_fill_em_kernel(float *ratio, const float *em, int W, int B, int Bp)
{
    const int w = blockIdx.x, a = blockIdx.y;
    const float e = em[a];
    float *r = ratio + ((int64_t)a * W + w) * Bp;
    for (int b = threadIdx.x; b < B; b += 64) r[b] = e;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static SmGeom sm_geometry(int64_t A, int64_t B, int64_t H, int64_t W, int spl = 0)
{
    SmGeom g{};
    g.A = (int)A;
    g.B = (int)B;
    g.H = (int)H;
    g.W = (int)W;
    g.nsi = (int)((H - 1 + kSmT - 1) / kSmT);
    g.nsj = (int)((W - 1 + kSmT - 1) / kSmT);
    const double cx = (W - 1) * 0.5, cy = (H - 1) * 0.5;
    const double reach = std::sqrt(cx * cx + cy * cy) + 2.0;
    g.reach = (float)reach;
    // deepest sample beyond the clamp box, in bins of 8 (the last bin is open)
    g.nd = (int)std::ceil((reach - std::min(cx, cy) + 2.0) / 8.0) + 1;
    g.nch = (int)std::max<int64_t>(1, std::min<int64_t>(8, A / 16));
    // slices per lane: as many as the slices fill (B must be a multiple of a given spl)
    g.spl = spl > 0 ? spl : (B % 4 == 0 && B >= 192) ? 4 : (B % 2 == 0 && B >= 96) ? 2 : 1;
    g.S = 64 * g.spl;
    g.Gi = (int)((B + g.S - 1) / g.S);
    g.nbi = (int)((H - 1 + kSmTB - 1) / kSmTB);
    g.nbj = (int)((W - 1 + kSmTB - 1) / kSmTB);
    return g;
}

// Scratch regions in floats: block sums, block halos, edge windows, their sums.
static void sm_sizes(const SmGeom &g, size_t n[4])
{
    const int ns = std::max(g.nsi, g.nsj);
    n[0] = (size_t)g.Gi * g.nbi * kSmTB * g.nbj * kSmTB * g.S;
    n[1] = (size_t)g.Gi * g.nbi * g.nbj * kSmHaloB * g.S;
    n[2] = (size_t)g.Gi * 4 * ns * g.nd * g.nch * kSmEdgeVox * g.S;
    n[3] = (size_t)g.Gi * 4 * ns * kSmEdgeVox * g.S;
}

static size_t sm_align(size_t floats) { return (floats * sizeof(float) + 255) / 256 * 256; }

size_t sm_scratch_bytes(int64_t A, int64_t B, int64_t H, int64_t W, int spl)
{
    const SmGeom g = sm_geometry(A, B, H, W, spl);
    size_t n[4];
    sm_sizes(g, n);
    return sm_align(n[0]) + sm_align(n[1]) + sm_align(n[2]) + sm_align(n[3]);
}

int launch_sm_transpose_pitched(const float *src, float *dst, int64_t N, int64_t B, int64_t L, int64_t Bp,
                                hipStream_t s)
{
    if (N == 0 || B == 0 || L == 0) return FL_OK;
    FL_REQUIRE(N < 65536 && B <= Bp && Bp < (1 << 24) && (L + 63) / 64 < (int64_t)UINT32_MAX,
               "slice-major copy: shape too large");
    sm_transpose_kernel<<<dim3((uint32_t)((L + 63) / 64), (uint32_t)((B + 63) / 64), (uint32_t)N), dim3(256), 0,
                          s>>>(src, dst, (int)B, L, (int)Bp);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_sm_transpose(const float *src, float *dst, int64_t N, int64_t B, int64_t L, hipStream_t s)
{
    return launch_sm_transpose_pitched(src, dst, N, B, L, B, s);
}

int launch_sm_fill_em(float *ratio, const float *em, int64_t A, int64_t W, int64_t B, int64_t Bp, hipStream_t s)
{
    if (A == 0 || W == 0 || B == 0) return FL_OK;
    FL_REQUIRE(A < 65536 && W < ((int64_t)1 << 31) && B <= Bp && Bp < (1 << 24), "em fill: shape too large");
    sm_fill_em_kernel<<<dim3((uint32_t)W, (uint32_t)A), dim3(64), 0, s>>>(ratio, em, (int)W, (int)B, (int)Bp);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

// Launch records of one problem: the geometry and the scratch regions.
static SmArgs sm_args(const SmGeom &g, const float *atten_sm, const float *ratio, const float *em,
                      const float *theta, void *scratch)
{
    size_t n[4];
    sm_sizes(g, n);
    char *ws = (char *)scratch;
    SmArgs p{};
    p.g = g;
    p.atten = atten_sm;
    p.ratio = ratio;
    p.em = em;
    p.theta = theta;
    p.own = (float *)ws;
    p.halo = (float *)(ws + sm_align(n[0]));
    p.edge = (float *)(ws + sm_align(n[0]) + sm_align(n[1]));
    p.edge_sum = (float *)(ws + sm_align(n[0]) + sm_align(n[1]) + sm_align(n[2]));
    return p;
}

static SmFin sm_fin(const SmArgs &p, float *out, int mode, const float *C, const float *sens, float beta,
                    float delta)
{
    SmFin f{};
    f.g = p.g;
    f.own = p.own;
    f.halo = p.halo;
    f.edge_sum = p.edge_sum;
    f.out = out;
    f.mode = mode;
    f.C = C;
    f.sens = sens;
    f.beta = beta;
    f.delta = delta;
    return f;
}

// Workgroups of each launch of one problem.
static int64_t sm_interior_wgs(const SmGeom &g) { return (int64_t)g.Gi * g.nbi * g.nbj; }
static int64_t sm_edge_wgs(const SmGeom &g) { return (int64_t)g.Gi * sm_edge_items(g); }
// Fused interior + edge launch: quads of interior blocks, then the edge items,
// each padded to a multiple of the XCD count (in a grouped launch every job
// then starts on XCD 0, as its own launch would).
static int64_t sm_fused_wgs(const SmGeom &g)
{
    const int64_t x = kNumXcd;
    return ((sm_interior_wgs(g) + kSmEW - 1) / kSmEW + x - 1) / x * x + (sm_edge_wgs(g) + x - 1) / x * x;
}
static int64_t sm_reduce_wgs(const SmGeom &g) { return (int64_t)g.Gi * 4 * std::max(g.nsi, g.nsj) * kSmEdgeVox; }
static int64_t sm_finalize_wgs(const SmGeom &g) { return (int64_t)((g.W + 63) / 64) * g.H * ((g.B + 63) / 64); }

template <int SPL>
static void sm_launch(const SmArgs &p, uint32_t n_fused, uint32_t n_red, hipStream_t stream)
{
    if (p.ratio)
        sm_fused_kernel<SPL, true><<<dim3(n_fused), dim3(64 * kSmEW), 0, stream>>>(p);
    else
        sm_fused_kernel<SPL, false><<<dim3(n_fused), dim3(64 * kSmEW), 0, stream>>>(p);
    sm_edge_reduce_kernel<SPL><<<dim3(n_red), dim3(64), 0, stream>>>(p);
}

int launch_sm_backward(const float *atten_sm, const float *ratio_sm, const float *em, const float *theta,
                       float *out, int64_t A, int64_t B, int64_t H, int64_t W, int mode, const float *C,
                       const float *sens, float beta, float delta, void *scratch, size_t scratch_bytes,
                       hipStream_t stream)
{
    const SmGeom g = sm_geometry(A, B, H, W);
    FL_REQUIRE(scratch_bytes >= sm_scratch_bytes(A, B, H, W) && scratch, "slice-major back-projector: scratch too small");
    FL_REQUIRE(H * W * B < ((int64_t)1 << 29), "slice-major back-projector: H*W*B must be below 2^29");
    const SmArgs p = sm_args(g, atten_sm, ratio_sm, em, theta, scratch);
    const int64_t n_fused = sm_fused_wgs(g);
    const int64_t n_red = sm_reduce_wgs(g);
    FL_REQUIRE(n_fused < ((int64_t)1 << 24) && n_red < ((int64_t)1 << 26), "slice-major back-projector: grid too large");
    if (g.spl == 4)
        sm_launch<4>(p, (uint32_t)n_fused, (uint32_t)n_red, stream);
    else if (g.spl == 2)
        sm_launch<2>(p, (uint32_t)n_fused, (uint32_t)n_red, stream);
    else
        sm_launch<1>(p, (uint32_t)n_fused, (uint32_t)n_red, stream);
    FL_LAUNCH_CHECK();
    const SmFin f = sm_fin(p, out, mode, C, sens, beta, delta);
    sm_finalize_kernel<<<dim3((uint32_t)((W + 63) / 64), (uint32_t)H, (uint32_t)((B + 63) / 64)), dim3(64), 0,
                         stream>>>(f);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

// ---------------------------------------------------------------------------
// Grouped launches (batched MLEM)
// ---------------------------------------------------------------------------
static size_t tb_align(size_t b) { return (b + 255) / 256 * 256; }

// Table image: SmArgs[n] (grouped by slices-per-lane class), two begin arrays
// per class (fused edge + interior, reduce), then n_var SmFin[n] and the
// finalize begin array.
size_t sm_group_table_bytes(int n)
{
    if (n < 1) return 0;
    size_t b = tb_align(n * sizeof(SmArgs));
    b += 3 * 2 * tb_align((size_t)(n + 1) * sizeof(uint32_t));
    b += kSmGroupVariants * tb_align(n * sizeof(SmFin));
    b += tb_align((size_t)(n + 1) * sizeof(uint32_t));
    return b;
}

int sm_group_build(const SmPack *packs, int n, const SmFinSpec *fin, int n_var, char *host, SmGroup *out)
{
    FL_REQUIRE(n >= 1 && n_var >= 1 && n_var <= kSmGroupVariants && packs && fin && host && out,
               "slice-major group: bad arguments");
    std::vector<SmGeom> geo(n);
    for (int i = 0; i < n; ++i) {
        const SmPack &q = packs[i];
        FL_REQUIRE(q.A >= 1 && q.B >= 1 && q.H >= 2 && q.W >= 2 && q.A < 65536 && q.H * q.W * q.B < ((int64_t)1 << 29),
                   "slice-major group: a problem's shape is out of range (H*W*B must be below 2^29)");
        FL_REQUIRE(q.atten_sm && q.ratio && q.em && q.theta && q.scratch, "slice-major group: null pointer");
        FL_REQUIRE((q.spl == 1 || q.spl == 2 || q.spl == 4) && q.B % q.spl == 0,
                   "slice-major group: slices per lane must be 1, 2 or 4 and divide B");
        geo[i] = sm_geometry(q.A, q.B, q.H, q.W, q.spl);
    }
    SmGroup G;
    G.n = n;
    G.n_var = n_var;
    size_t o = tb_align(n * sizeof(SmArgs));
    const size_t begin_bytes = tb_align((size_t)(n + 1) * sizeof(uint32_t));
    SmArgs *jobs = (SmArgs *)host;
    int k = 0;
    const int spls[3] = {4, 2, 1};
    for (int c = 0; c < 3; ++c) {
        // this class's problems, most angles first (their workgroups run longest)
        std::vector<int> idx;
        for (int i = 0; i < n; ++i)
            if (geo[i].spl == spls[c]) idx.push_back(i);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return geo[a].A > geo[b].A; });
        SmGroup::Class &cl = G.cls[c];
        cl.spl = spls[c];
        cl.n = (int)idx.size();
        cl.jobs = (size_t)k * sizeof(SmArgs);
        cl.b_fused = o;
        cl.b_red = o + begin_bytes;
        uint32_t *bf = (uint32_t *)(host + cl.b_fused), *br = (uint32_t *)(host + cl.b_red);
        int64_t tf = 0, tr = 0;
        for (size_t m = 0; m < idx.size(); ++m) {
            const SmPack &q = packs[idx[m]];
            jobs[k + m] = sm_args(geo[idx[m]], q.atten_sm, q.ratio, q.em, q.theta, q.scratch);
            bf[m] = (uint32_t)tf;
            br[m] = (uint32_t)tr;
            tf += sm_fused_wgs(geo[idx[m]]);
            tr += sm_reduce_wgs(geo[idx[m]]);
        }
        // workgroups x threads must stay below 2^32
        FL_REQUIRE(tf < ((int64_t)1 << 24) && tr < ((int64_t)1 << 26),
                   "slice-major group: too many workgroups for one launch");
        bf[idx.size()] = (uint32_t)tf;
        br[idx.size()] = (uint32_t)tr;
        cl.t_fused = (uint32_t)tf;
        cl.t_red = (uint32_t)tr;
        k += (int)idx.size();
        o += 2 * begin_bytes;
    }
    // finalize variants in pack order
    for (int v = 0; v < n_var; ++v) {
        G.fin[v] = o;
        SmFin *fv = (SmFin *)(host + o);
        for (int i = 0; i < n; ++i) {
            const SmPack &q = packs[i];
            const SmFinSpec &f = fin[(size_t)v * n + i];
            FL_REQUIRE(f.out != nullptr, "slice-major group: null output");
            const SmArgs p = sm_args(geo[i], q.atten_sm, q.ratio, q.em, q.theta, q.scratch);
            fv[i] = sm_fin(p, f.out, f.mode, f.C, f.sens, f.beta, f.delta);
        }
        o += tb_align(n * sizeof(SmFin));
    }
    o += (size_t)(kSmGroupVariants - n_var) * tb_align(n * sizeof(SmFin));
    G.b_fin = o;
    uint32_t *bf = (uint32_t *)(host + o);
    int64_t tf = 0;
    for (int i = 0; i < n; ++i) {
        bf[i] = (uint32_t)tf;
        tf += sm_finalize_wgs(geo[i]);
    }
    FL_REQUIRE(tf < ((int64_t)1 << 26), "slice-major group: too many workgroups for one launch");
    bf[n] = (uint32_t)tf;
    G.t_fin = (uint32_t)tf;
    *out = G;
    return FL_OK;
}

template <int SPL>
static void sm_group_class(const SmGroup &G, const SmGroup::Class &c, hipStream_t s)
{
    const SmArgs *jobs = (const SmArgs *)(G.dev + c.jobs);
    const uint32_t *bf = (const uint32_t *)(G.dev + c.b_fused), *br = (const uint32_t *)(G.dev + c.b_red);
    if (c.t_fused) sm_fused_group_kernel<SPL><<<dim3(c.t_fused), dim3(64 * kSmEW), 0, s>>>(jobs, bf, c.n);
    if (c.t_red) sm_edge_reduce_group_kernel<SPL><<<dim3(c.t_red), dim3(64), 0, s>>>(jobs, br, c.n);
}

int launch_sm_group(const SmGroup &G, int variant, hipStream_t s)
{
    FL_REQUIRE(G.dev != nullptr && variant >= 0 && variant < G.n_var, "slice-major group: bad variant");
    for (int c = 0; c < 3; ++c) {
        const SmGroup::Class &cl = G.cls[c];
        if (cl.n == 0) continue;
        if (cl.spl == 4)
            sm_group_class<4>(G, cl, s);
        else if (cl.spl == 2)
            sm_group_class<2>(G, cl, s);
        else
            sm_group_class<1>(G, cl, s);
        FL_LAUNCH_CHECK();
    }
    sm_finalize_group_kernel<<<dim3(G.t_fin), dim3(64), 0, s>>>((const SmFin *)(G.dev + G.fin[variant]),
                                                                 (const uint32_t *)(G.dev + G.b_fin), G.n);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

}  // namespace fl
