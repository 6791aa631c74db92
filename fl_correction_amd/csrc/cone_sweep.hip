// Direction-sampled detector cone, all directions in one sweep over x planes.
//
// Mask frame (Z, X, Y), cone along +X.  For direction d with slopes
// (pz_d, py_d) / Q (every direction's denominator divides Q):
//   J_d(z, x, y) = S_d(z, x, y) + J_d(z + pz_d, x + Q, y + py_d),
//   S_d(z, x, y) = sum_{t=1..Q} bilin0(mu[., x+t, .], z + t*pz_d/Q, y + t*py_d/Q)
// (the ray sum split into blocks of Q steps; every Q steps the ray sits on a
// voxel, so the recursion is exact), and T = sum_d w_d J_d.  The adjoint runs
// the transposed stencil from the near end: J_d(v) = S^T_d(v) + J_d(v - D_d).
//
// The x axis is cut into blocks of Q planes.  A plane's J needs only the plane
// Q further along the ray, which lies in the previous block at the same
// position, so the planes of one block are independent: one launch per block.
// A workgroup owns a tile of TZ x 64 voxels in (z, y) for R of the block's
// planes (R = Q unless the stencil's halo makes the staging too large): it
// stages the mu planes its taps reach (its planes shifted by 1..Q, with the
// stencil's halo in z and y) in LDS once, then computes every direction's
// stencil from LDS: per direction, each (z, y) column its taps touch is read
// once as a window of R + Q - 1 planes into registers, and the column's weights
// (one per staged plane, read through the scalar cache) are applied to all R
// planes from there.  Each direction's J of the block goes to a
// ring in device memory, read by the next block's launch at the shifted
// position; T accumulates in registers over the directions in order and is
// written once.  The chain form (cone_kernel in atten.hip) instead reads and
// rewrites T once per direction.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fl {

#include "cone_fixed.inc"  // tile geometry of the compiled-in direction grids

constexpr int kSweepMaxDir = 64;
constexpr int kSweepLanes = 64;
constexpr uint32_t kSweepOOB = 0x80000000u;  // a byte offset past every buffer (reads 0)

struct SweepSetup {
    int ndir, Q, reverse;
    int cols, NP;          // LDS tile: row length, floats per (z, y) position (the planes, odd)
    int dzmin, dymin;      // lowest tap offsets in z and y
    int pz[kSweepMaxDir], py[kSweepMaxDir];  // slopes scaled to Q
    float w[kSweepMaxDir];
};

// Per direction: the columns of the tile its taps reach, at most 4Q, each as
// Q + 1 words at d * 4Q * (Q + 1): the LDS offset of the column's staged plane
// 0, then the bilinear weight of the tap on each staged plane (0 where the
// direction has none); their count, the shift of the recursion (signed for the
// mode) and the direction's weight.
struct SweepTable {
    int *cols;
    int *count;
    int2 *shift;
    float *weight;
};

static int64_t floordiv64(int64_t a, int64_t b)
{
    int64_t q = a / b, r = a % b;
    return (r != 0 && ((r < 0) != (b < 0))) ? q - 1 : q;
}

__device__ __forceinline__ int floordiv_dev(int a, int b)
{
    int q = a / b, r = a % b;
    return (r != 0 && ((r < 0) != (b < 0))) ? q - 1 : q;
}

// One thread per direction: the bilinear taps of its Q samples, weights rounded
// from their exact rational values, zero weights dropped, gathered by column.
__global__ void cone_sweep_setup_kernel(SweepSetup s, SweepTable tb)
{
    // offsets of the columns found so far (column-major: neighbouring lanes in
    // different banks), searched here rather than in the table just written
    __shared__ int found[4 * 16][kSweepMaxDir];
    const int d = threadIdx.x;
    if (d >= s.ndir) return;
    const int sg = s.reverse ? -1 : 1;
    int n = 0;
    // columns of this direction: (LDS offset of staged plane 0, weight by staged plane)
    int *out = tb.cols + d * 4 * s.Q * (s.Q + 1);
    for (int t = 1; t <= s.Q; ++t) {
        const int nz = t * s.pz[d], ny = t * s.py[d];
        const int z0 = floordiv_dev(nz, s.Q), y0 = floordiv_dev(ny, s.Q);
        const double fz = (double)(nz - z0 * s.Q) / s.Q, fy = (double)(ny - y0 * s.Q) / s.Q;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const double wz = a ? fz : 1.0 - fz, wy = b ? fy : 1.0 - fy;
                if (wz == 0.0 || wy == 0.0) continue;
                const int dz = sg * (z0 + a), dy = sg * (y0 + b);
                // staged plane of x + sg*t relative to the voxel's plane
                const int pl = s.reverse ? s.Q - t : t - 1;
                const int off = ((dz - s.dzmin) * s.cols + (dy - s.dymin)) * s.NP;
                int c = 0;
                while (c < n && found[c][d] != off) ++c;
                if (c == n) {
                    found[c][d] = off;
                    out[c * (s.Q + 1)] = off;
                    for (int j = 0; j < s.Q; ++j) out[c * (s.Q + 1) + 1 + j] = 0;
                    ++n;
                }
                out[c * (s.Q + 1) + 1 + pl] = __float_as_int((float)(wz * wy));
            }
    }
    tb.count[d] = n;
    tb.shift[d] = make_int2(sg * s.pz[d], sg * s.py[d]);
    tb.weight[d] = s.w[d];
}

struct SweepArgs {
    const float *mu;
    float *T;
    const float *ring_in;
    float *ring_out;
    SweepTable tb;
    int K, Z, X, Y, Q, ndir;
    int xs, nx, pos0;  // the block's first plane, its plane count, that plane's position
    int first, reverse;
    int TZ, rows, cols, PS, NP, planes, dzmin, dymin;
};

// FM > 0: the direction grid {-FM..FM}^2 / 8 of cone_fixed.inc, whose stencil is
// compiled in (Q = R = 8, the tile geometry the plan makes for it).
template <int Q, int R, int FM = 0>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(FM ? 4 : 1))) void cone_sweep_kernel(
    SweepArgs p)
{
    extern __shared__ float win[];
    const int k = blockIdx.z;
    const int zt = blockIdx.y * p.TZ, yt = blockIdx.x * kSweepLanes;
    const int lane = threadIdx.x & (kSweepLanes - 1), zl = threadIdx.x / kSweepLanes;
    const int64_t ZXY = (int64_t)p.Z * p.X * p.Y;
    const float *muk = p.mu + k * ZXY;

    // ---- stage the planes the launch's taps reach ---------------------------
    // forward: x = xs+1 .. xs+R-1+Q; adjoint: x = xs-Q .. xs+R-2 (R + Q - 1 planes)
    const int xfirst = p.reverse ? p.xs - Q : p.xs + 1;
    // the per-direction records, after the tile in LDS: their reads then wait
    // on the LDS counter only, not behind the ring's loads and stores
    int4 *dirs = reinterpret_cast<int4 *>(win + ((p.NP * p.PS + 3) & ~3));  // (count, shift z, shift y, weight bits)
    for (int d = threadIdx.x; d < p.ndir; d += blockDim.x) {
        const int2 sh = p.tb.shift[d];
        dirs[d] = make_int4(p.tb.count[d], sh.x, sh.y, __float_as_int(p.tb.weight[d]));
    }
    // buffer loads: an offset past the line's volume reads 0 (no branches)
    const __amdgpu_buffer_rsrc_t rmu = make_rsrc(muk, (int)ZXY);
    // the planes of one (z, y) position are neighbours in LDS (NP floats, odd, so
    // that 64 lanes at consecutive y hit 64 different banks): a tap's R planes are
    // then read as pairs of words
    const int total = p.planes * p.PS;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int pl = idx / p.PS, rem = idx - pl * p.PS;
        const int r = rem / p.cols, c = rem - r * p.cols;
        const int x = xfirst + pl, z = zt + p.dzmin + r, y = yt + p.dymin + c;
        const bool in = (unsigned)x < (unsigned)p.X && (unsigned)z < (unsigned)p.Z && (unsigned)y < (unsigned)p.Y;
        win[rem * p.NP + pl] = load_f32(rmu, in ? (uint32_t)(((int64_t)z * p.X + x) * p.Y + y) * 4u : kSweepOOB);
    }
    __syncthreads();

    const int z = zt + zl, y = yt + lane;
    const bool act = z < p.Z && y < p.Y;
    const float *wbase = win + (zl * p.cols + lane) * p.NP;
    const int64_t ring_dir = (int64_t)Q * p.Z * p.Y;
    float Tacc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) Tacc[r] = 0.0f;

    // the recursion's other end of direction d: the previous block, same
    // position, shifted; loaded one direction ahead of its use
    auto load_jp = [&](int d, float *jp) {
        const int4 di = dirs[d];
        const int2 sh = make_int2(di.y, di.z);
        const __amdgpu_buffer_rsrc_t rin = make_rsrc(p.ring_in + ((int64_t)k * p.ndir + d) * ring_dir, (int)ring_dir);
        const int zz = z + sh.x, yy = y + sh.y;
        const bool in_zy = act && (unsigned)zz < (unsigned)p.Z && (unsigned)yy < (unsigned)p.Y && !p.first;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int x = p.xs + r;
            const int xx = p.reverse ? x - Q : x + Q;
            const bool ok = r < p.nx && in_zy && (unsigned)xx < (unsigned)p.X;
            jp[r] = load_f32(rin, ok ? (uint32_t)(((p.pos0 + r) * p.Z + zz) * p.Y + yy) * 4u : kSweepOOB);
        }
    };
    // the end of direction d: J = S + J_prev to the ring, and T
    auto finish = [&](int d, const float *S, const float *jp) {
        const float wd = __int_as_float(dirs[d].w);
        float *rout = p.ring_out + ((int64_t)k * p.ndir + d) * ring_dir;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r < p.nx) {
                const float J = S[r] + jp[r];
                if (act) rout[((int64_t)(p.pos0 + r) * p.Z + z) * p.Y + y] = J;
                Tacc[r] = d == 0 ? wd * J : fmaf(wd, J, Tacc[r]);
            }
        }
    };
    if constexpr (FM == 0) {
        float jp[R], jn[R];
        load_jp(0, jp);
        for (int d = 0; d < p.ndir; ++d) {
            const int4 di = dirs[d];
            const float wd = __int_as_float(di.w);
            const int cnt = di.x;
            using ctab_t = const __attribute__((address_space(4))) int *;
            const ctab_t ct = (ctab_t)(p.tb.cols) + d * 4 * Q * (Q + 1);
            float *rout = p.ring_out + ((int64_t)k * p.ndir + d) * ring_dir;
            if (d + 1 < p.ndir) load_jp(d + 1, jn);
            // every staged plane r < R is computed (planes past nx are in the tile
            // and are not written): no branch between the LDS reads
            float S[R];
    #pragma unroll
            for (int r = 0; r < R; ++r) S[r] = 0.0f;
            for (int c = 0; c < cnt; ++c) {
                const ctab_t e = ct + c * (Q + 1);
                const float *src = wbase + e[0];
                float col[R + Q - 1];
    #pragma unroll
                for (int k = 0; k < R + Q - 1; ++k) col[k] = src[k];
    #pragma unroll
                for (int j = 0; j < Q; ++j) {
                    const float w = __int_as_float(e[1 + j]);
    #pragma unroll
                    for (int r = 0; r < R; ++r) S[r] = fmaf(w, col[j + r], S[r]);
                }
            }
    #pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r < p.nx) {
                    const float J = S[r] + jp[r];
                    if (act) rout[((int64_t)(p.pos0 + r) * p.Z + z) * p.Y + y] = J;
                    Tacc[r] = d == 0 ? wd * J : fmaf(wd, J, Tacc[r]);
                }
            }
    #pragma unroll
            for (int r = 0; r < R; ++r) jp[r] = jn[r];
        }
    } else {
        // Groups of up to G directions with one pz: each column the group's taps
        // touch is read once (a window of R + Q - 1 planes) and its taps, with
        // literal weights, add into their directions' plane sums.
        constexpr int G = 3;  // at most G directions per group (gen_cone_fixed.py GMAX)
#define CF_GROUP_BEGIN(d0)                                      \
    {                                                           \
        float S[G][R], c[R + Q - 1];                            \
        _Pragma("unroll") for (int g = 0; g < G; ++g)           \
            _Pragma("unroll") for (int r = 0; r < R; ++r) S[g][r] = 0.0f;
#define CF_COL(off) _Pragma("unroll") for (int i = 0; i < R + Q - 1; ++i) c[i] = wbase[(off) + i];
#define CF_TAP(g, j, w) _Pragma("unroll") for (int r = 0; r < R; ++r) S[g][r] = fmaf((w), c[(j) + r], S[g][r]);
#define CF_GROUP_END(d0, n)                                     \
        _Pragma("unroll") for (int g = 0; g < (n); ++g)         \
        {                                                       \
            float jg[R];                                        \
            load_jp((d0) + g, jg);                              \
            finish((d0) + g, S[g], jg);                         \
        }                                                       \
    }
        if constexpr (FM == 1) {
            if (!p.reverse) {
#define CF_M1_FWD_BODY
#include "cone_fixed.inc"
#undef CF_M1_FWD_BODY
            } else {
#define CF_M1_ADJ_BODY
#include "cone_fixed.inc"
#undef CF_M1_ADJ_BODY
            }
        } else {
            if (!p.reverse) {
#define CF_M2_FWD_BODY
#include "cone_fixed.inc"
#undef CF_M2_FWD_BODY
            } else {
#define CF_M2_ADJ_BODY
#include "cone_fixed.inc"
#undef CF_M2_ADJ_BODY
            }
        }
#undef CF_GROUP_BEGIN
#undef CF_COL
#undef CF_TAP
#undef CF_GROUP_END
    }
    if (act) {
        float *Tk = p.T + k * ZXY;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r < p.nx) Tk[((int64_t)z * p.X + p.xs + r) * p.Y + y] = Tacc[r];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct SweepPlan {
    int Q = 0, R = 0, TZ = 0, rows = 0, cols = 0, PS = 0, NP = 0, planes = 0, dzmin = 0, dymin = 0;
    std::vector<int> pz, py;
};

static size_t sweep_align(size_t bytes) { return (bytes + 255) / 256 * 256; }

// LDS of one workgroup: the tile's planes, then the per-direction records.
static size_t sweep_lds_bytes(const SweepPlan &pl, int ndir)
{
    return (size_t)((pl.NP * pl.PS + 3) & ~3) * sizeof(float) + (size_t)ndir * sizeof(int4);
}

// Geometry of the sweep for a direction set; Q = 0 when it does not apply.
static SweepPlan sweep_plan(const int32_t *pz, const int32_t *py, const int32_t *q, int64_t ndir, int reverse)
{
    SweepPlan pl;
    if (ndir < 1 || ndir > kSweepMaxDir) return pl;
    int Q = 0;
    for (int64_t d = 0; d < ndir; ++d) Q = std::max(Q, (int)q[d]);
    if (Q != 4 && Q != 8 && Q != 16) return pl;
    int dzmin = INT32_MAX, dzmax = INT32_MIN, dymin = INT32_MAX, dymax = INT32_MIN;
    const int sg = reverse ? -1 : 1;
    for (int64_t d = 0; d < ndir; ++d) {
        if (q[d] < 1 || Q % q[d] != 0) return SweepPlan{};
        const int64_t m = Q / q[d];
        if (std::llabs((int64_t)pz[d] * m) > 4 * Q || std::llabs((int64_t)py[d] * m) > 4 * Q) return SweepPlan{};
        pl.pz.push_back((int)(pz[d] * m));
        pl.py.push_back((int)(py[d] * m));
        for (int t = 1; t <= Q; ++t) {
            const int64_t z0 = floordiv64((int64_t)t * pl.pz.back(), Q), y0 = floordiv64((int64_t)t * pl.py.back(), Q);
            // both bilinear neighbours bound the reach (a zero weight costs nothing)
            for (int a = 0; a < 2; ++a) {
                dzmin = std::min(dzmin, (int)(sg * (z0 + a)));
                dzmax = std::max(dzmax, (int)(sg * (z0 + a)));
                dymin = std::min(dymin, (int)(sg * (y0 + a)));
                dymax = std::max(dymax, (int)(sg * (y0 + a)));
            }
        }
    }
    pl.cols = kSweepLanes + (dymax - dymin);
    // R planes of a block per launch (R + Q - 1 staged planes), TZ rows of voxels
    // per workgroup: the pair with the most voxels per staged value that fits
    double best = 0.0;
    for (int R = Q; R >= Q / 4; R /= 2)
        for (int TZ : {8, 4, 2, 1}) {
            const int rows = TZ + (dzmax - dzmin);
            const int planes = R + Q - 1;
            const size_t table = (size_t)ndir * sizeof(int4);
            if ((size_t)((planes | 1) * rows * pl.cols + 3) * sizeof(float) + table > 96 * 1024) continue;
            const double score = (double)TZ * R / ((double)planes * rows);
            if (score > best) {
                best = score;
                pl.TZ = TZ;
                pl.R = R;
                pl.rows = rows;
                pl.planes = planes;
            }
        }
    if (!pl.TZ) return SweepPlan{};
    pl.Q = Q;
    pl.PS = pl.rows * pl.cols;
    pl.NP = pl.planes | 1;
    pl.dzmin = dzmin;
    pl.dymin = dymin;
    return pl;
}

size_t cone_sweep_workspace_bytes(int64_t K, int64_t Z, int64_t X, int64_t Y, const int32_t *pz, const int32_t *py,
                                  const int32_t *q, int64_t ndir)
{
    (void)X;
    const SweepPlan pl = sweep_plan(pz, py, q, ndir, 0);
    if (!pl.Q) return 0;
    const size_t ring = sweep_align((size_t)K * ndir * pl.Q * Z * Y * sizeof(float));
    const size_t table = sweep_align((size_t)ndir * 4 * pl.Q * (pl.Q + 1) * sizeof(int)) + sweep_align(ndir * sizeof(int)) +
                         sweep_align(ndir * sizeof(int2)) + sweep_align(ndir * sizeof(float));
    return 2 * ring + table;
}

template <int Q, int R>
static const void *sweep_fn()
{
    return (const void *)cone_sweep_kernel<Q, R>;
}

template <int Q, int R, int FM = 0>
static void sweep_launch(const SweepArgs &a, dim3 grid, int threads, size_t lds, hipStream_t s)
{
    cone_sweep_kernel<Q, R, FM><<<grid, dim3(threads), lds, s>>>(a);
}

// 1 or 2: the plan's directions are the grid {-M..M}^2 / 8 of cone_fixed.inc,
// in its order, with the tile geometry compiled in; else 0.
static int fixed_grid(const SweepPlan &pl, int64_t ndir, int reverse)
{
    if (pl.Q != 8 || pl.R != 8 || pl.TZ != 8 || (ndir != 9 && ndir != 25)) return 0;
    const int M = ndir == 9 ? 1 : 2, side = 2 * M + 1;
    for (int64_t d = 0; d < ndir; ++d)
        if (pl.pz[d] != (int)(d / side) - M || pl.py[d] != (int)(d % side) - M) return 0;
    struct Geo {
        int dzmin, dymin, cols, rows, np;
    };
    const Geo g = M == 1 ? (reverse ? Geo{CF_M1_ADJ_DZMIN, CF_M1_ADJ_DYMIN, CF_M1_ADJ_COLS, CF_M1_ADJ_ROWS, CF_M1_ADJ_NP}
                                    : Geo{CF_M1_FWD_DZMIN, CF_M1_FWD_DYMIN, CF_M1_FWD_COLS, CF_M1_FWD_ROWS, CF_M1_FWD_NP})
                         : (reverse ? Geo{CF_M2_ADJ_DZMIN, CF_M2_ADJ_DYMIN, CF_M2_ADJ_COLS, CF_M2_ADJ_ROWS, CF_M2_ADJ_NP}
                                    : Geo{CF_M2_FWD_DZMIN, CF_M2_FWD_DYMIN, CF_M2_FWD_COLS, CF_M2_FWD_ROWS, CF_M2_FWD_NP});
    const bool same = pl.dzmin == g.dzmin && pl.dymin == g.dymin && pl.cols == g.cols && pl.rows == g.rows &&
                      pl.NP == g.np;
    return same ? M : 0;
}

// The kernel instance for (Q, R), R in {Q, Q/2, Q/4}.
template <int Q>
static void sweep_pick(int R, const void **fn, void (**launch)(const SweepArgs &, dim3, int, size_t, hipStream_t))
{
    if (R == Q) {
        *fn = sweep_fn<Q, Q>();
        *launch = sweep_launch<Q, Q>;
    } else if (R == Q / 2) {
        *fn = sweep_fn<Q, Q / 2>();
        *launch = sweep_launch<Q, Q / 2>;
    } else {
        *fn = sweep_fn<Q, Q / 4>();
        *launch = sweep_launch<Q, Q / 4>;
    }
}

int launch_cone_sweep(const float *mu, float *T, int64_t K, int64_t Z, int64_t X, int64_t Y, const int32_t *pz,
                      const int32_t *py, const int32_t *q, const float *weight, int64_t ndir, int reverse,
                      void *workspace, size_t workspace_bytes, hipStream_t s)
{
    const SweepPlan pl = sweep_plan(pz, py, q, ndir, reverse);
    if (!pl.Q) {
        set_error("cone sweep: unsupported direction set (denominators must divide one Q in {4, 8, 16}, "
                  "at most 64 directions, the stencil must fit the LDS tile)");
        return FL_ERR_UNSUPPORTED;
    }
    FL_REQUIRE(Z * X * Y < ((int64_t)1 << 29) && (int64_t)pl.Q * Z * Y < ((int64_t)1 << 29),
               "cone sweep: one line's volume must hold fewer than 2^29 voxels");
    FL_REQUIRE(K < 65536 && (Z + pl.TZ - 1) / pl.TZ < 65536, "cone sweep: grid too large");
    const size_t need = cone_sweep_workspace_bytes(K, Z, X, Y, pz, py, q, ndir);
    if (!workspace || workspace_bytes < need) {
        set_error("cone sweep: workspace too small: need " + std::to_string(need) + " bytes");
        return FL_ERR_WORKSPACE;
    }
    const size_t ring = sweep_align((size_t)K * ndir * pl.Q * Z * Y * sizeof(float));
    char *ws = (char *)workspace;
    float *rings[2] = {(float *)ws, (float *)(ws + ring)};
    char *tp = ws + 2 * ring;
    SweepTable tb;
    tb.cols = (int *)tp;
    tp += sweep_align((size_t)ndir * 4 * pl.Q * (pl.Q + 1) * sizeof(int));
    tb.count = (int *)tp;
    tp += sweep_align(ndir * sizeof(int));
    tb.shift = (int2 *)tp;
    tp += sweep_align(ndir * sizeof(int2));
    tb.weight = (float *)tp;

    SweepSetup su{};
    su.ndir = (int)ndir;
    su.Q = pl.Q;
    su.reverse = reverse;
    su.cols = pl.cols;
    su.NP = pl.NP;
    su.dzmin = pl.dzmin;
    su.dymin = pl.dymin;
    for (int64_t d = 0; d < ndir; ++d) {
        su.pz[d] = pl.pz[d];
        su.py[d] = pl.py[d];
        su.w[d] = weight[d];
    }
    cone_sweep_setup_kernel<<<1, kSweepMaxDir, 0, s>>>(su, tb);
    FL_LAUNCH_CHECK();

    const size_t lds = sweep_lds_bytes(pl, (int)ndir);
    const void *fn = nullptr;
    void (*launch)(const SweepArgs &, dim3, int, size_t, hipStream_t) = nullptr;
    const int fm = fixed_grid(pl, ndir, reverse);
    if (fm == 1) {
        fn = (const void *)cone_sweep_kernel<8, 8, 1>;
        launch = sweep_launch<8, 8, 1>;
    } else if (fm == 2) {
        fn = (const void *)cone_sweep_kernel<8, 8, 2>;
        launch = sweep_launch<8, 8, 2>;
    } else if (pl.Q == 4) {
        sweep_pick<4>(pl.R, &fn, &launch);
    } else if (pl.Q == 8) {
        sweep_pick<8>(pl.R, &fn, &launch);
    } else {
        sweep_pick<16>(pl.R, &fn, &launch);
    }
    FL_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 grid((unsigned)((Y + kSweepLanes - 1) / kSweepLanes), (unsigned)((Z + pl.TZ - 1) / pl.TZ), (unsigned)K);
    const int nb = (int)((X + pl.Q - 1) / pl.Q);
    for (int j = 0; j < nb; ++j) {
        // the block's planes [bx0, bx1), position 0 at bpos
        int64_t bx0, bx1, bpos0;
        if (reverse) {  // blocks from x = 0 upward: block j holds x in [jQ, (j+1)Q)
            bx0 = (int64_t)j * pl.Q;
            bx1 = std::min<int64_t>(bx0 + pl.Q, X);
            bpos0 = 0;
        } else {  // blocks from the far end: block j holds x in [X-(j+1)Q, X-jQ)
            const int64_t xb = X - (int64_t)(j + 1) * pl.Q;
            bx0 = std::max<int64_t>(0, xb);
            bx1 = X - (int64_t)j * pl.Q;
            bpos0 = bx0 - xb;
        }
        for (int64_t cx = bx0; cx < bx1; cx += pl.R) {
        SweepArgs a{};
        a.mu = mu;
        a.T = T;
        a.ring_in = rings[(j + 1) & 1];
        a.ring_out = rings[j & 1];
        a.tb = tb;
        a.K = (int)K;
        a.Z = (int)Z;
        a.X = (int)X;
        a.Y = (int)Y;
        a.Q = pl.Q;
        a.ndir = (int)ndir;
        a.xs = (int)cx;
        a.nx = (int)std::min<int64_t>(pl.R, bx1 - cx);
        a.pos0 = (int)(bpos0 + (cx - bx0));
        a.first = j == 0;
        a.reverse = reverse;
        a.TZ = pl.TZ;
        a.rows = pl.rows;
        a.cols = pl.cols;
        a.PS = pl.PS;
        a.NP = pl.NP;
        a.planes = pl.planes;
        a.dzmin = pl.dzmin;
        a.dymin = pl.dymin;
        launch(a, grid, kSweepLanes * pl.TZ, lds, s);
        FL_LAUNCH_CHECK();
        }
    }
    return FL_OK;
}

}  // namespace fl
