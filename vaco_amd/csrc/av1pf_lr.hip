// av1pf_lr.hip - loop restoration (AV1 7.17) for gfx950.
//
// Reference semantics (matteopolak/vaco, crates/codec/vaco-codec-av1/src/
// restoration.rs): restore_plane 178-207, Source 209-245, wiener 259-283,
// box_ab 314-339, guided_pass 341-371, self_guided 373-384.
//
// One workgroup of 512 threads per tile: at most 64x64 output samples
// inside one restoration stripe and one restoration unit, so the unit type
// is uniform over the workgroup. The tile's source window (Source::get
// semantics: CDEF output inside the stripe, pre-CDEF rows across its
// borders, edge clamping) is staged in LDS once. Wiener runs as a horizontal
// pass into LDS followed by a vertical pass; self-guided builds column sums
// of x and x^2 in LDS, then box sums per centre, then the A/B grids, then
// the 3x3 weighted combination.
#include "av1pf_common.h"

namespace av1pf {

constexpr int kLrThreads = 512;
constexpr int kS = 72;   // staged row stride (samples), 70 used
constexpr int kG = 66;   // A/B grid row stride, 66 used

// SGR_PARAMS, restoration.rs:286-303: {r0, e0, r1, e1}.
__constant__ int kSgr[16][4] = {
    {2, 12, 1, 4},  {2, 15, 1, 6},  {2, 18, 1, 8},  {2, 21, 1, 9},
    {2, 24, 1, 10}, {2, 29, 1, 11}, {2, 36, 1, 12}, {2, 45, 1, 13},
    {2, 56, 1, 14}, {2, 68, 1, 15}, {0, 0, 1, 5},   {0, 0, 1, 8},
    {0, 0, 1, 11},  {0, 0, 1, 14},  {2, 30, 0, 0},  {2, 75, 0, 0},
};

// A/B of one box from its sums (restoration.rs:314-339). The products fit
// their types for 8/10/12-bit samples: sumsq <= 25 * 4095^2 < 2^31;
// p * scale and (256 - a2) * sum * recip are formed in 64 bits.
template <int N>
__device__ __forceinline__ void ab_of(int sum, int sumsq, int bd, uint32_t scale,
                                      const uint16_t *xbx1, int &a2, int &b2)
{
    constexpr uint32_t recip = ((1u << 12) + N / 2) / N;
    const int a = round2i(sumsq, 2 * (bd - 8));
    const int d = round2i(sum, bd - 8);
    const int p = max(a * N - d * d, 0);
    const uint64_t z = ((uint64_t)(uint32_t)p * scale + (1u << 19)) >> 20;
    a2 = z >= 255 ? 256 : xbx1[(int)z];
    b2 = (int)(((uint64_t)(uint32_t)(256 - a2) * (uint32_t)sum * recip + (1u << 11)) >> 12);
}

// Column sums over 2R+1 staged rows, then box sums and A/B for every centre
// (y0-1+cr, x0-1+cc), cr < th+2, cc < tw+2. Pass 0 (R = 2) needs odd rows only.
template <int R>
__device__ __forceinline__ void sgr_grid(const uint16_t *S, uint16_t *VS, int *VQ, int *A, int *B,
                                         int th, int tw, int y0, int bd, int eps,
                                         const uint16_t *xbx1)
{
    constexpr int N = (2 * R + 1) * (2 * R + 1);
    const uint32_t n2e = (uint32_t)(N * N * eps);
    const uint32_t scale = ((1u << 20) + n2e / 2) / n2e;
    const int crows = th + 2, ccols = tw + 2, scols = tw + 6;
    const int tid = threadIdx.x;
    for (int e = tid; e < crows * kS; e += kLrThreads) {
        const int cr = e / kS, c = e - cr * kS;
        if (c >= scols || (R == 2 && !((y0 - 1 + cr) & 1)))
            continue;
        int s = 0, q = 0;
#pragma unroll
        for (int dy = -R; dy <= R; dy++) {
            const int v = S[(cr + 2 + dy) * kS + c];
            s += v;
            q += v * v;
        }
        VS[cr * kS + c] = (uint16_t)s;
        VQ[cr * kS + c] = q;
    }
    __syncthreads();
    for (int e = tid; e < crows * kS; e += kLrThreads) {
        const int cr = e / kS, cc = e - cr * kS;
        if (cc >= ccols || (R == 2 && !((y0 - 1 + cr) & 1)))
            continue;
        int s = 0, q = 0;
#pragma unroll
        for (int dx = -R; dx <= R; dx++) {
            s += VS[cr * kS + cc + 2 + dx];
            q += VQ[cr * kS + cc + 2 + dx];
        }
        int a2, b2;
        ab_of<N>(s, q, bd, scale, xbx1, a2, b2);
        A[cr * kG + cc] = a2;
        B[cr * kG + cc] = b2;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kLrThreads) void k_lr(const LrPlane *__restrict__ planes, int n_planes)
{
    __shared__ __align__(16) uint16_t S[70 * kS];
    __shared__ __align__(16) int M[66 * kS + 66 * kS / 2];  // Wiener rows | VS (u16) + VQ
    __shared__ __align__(16) int AB[2 * kG * kG];           // A | B grids
    __shared__ uint16_t xbx1[256];

    const LrPlane &P = find_item(planes, n_planes, blockIdx.x);
    const uint32_t local = blockIdx.x - P.block_begin;
    const uint32_t rsi = local / P.n_cols;
    const LrRowSeg rs = P.rows[rsi];
    const LrColSeg cs = P.cols[local - rsi * P.n_cols];
    LrUnit u;
    u.kind = 0;
    if (P.units)
        u = P.units[rs.unit_row * P.unit_cols + cs.unit_col];
    const int tid = threadIdx.x;
    const int W = P.width, H = P.height;
    const int bd = P.bit_depth, maxv = (1 << bd) - 1;
    const int tw = cs.w, th = rs.h;
    const int x0 = cs.x0, y0 = rs.y0;

    if (u.kind != kUnitWiener && u.kind != kUnitSelfGuided) {
        // RestorationUnit::None (restoration.rs:194) and copy_visible_plane
        // (decode.rs:673-681): the CDEF output itself.
        bool bad = false;
#pragma unroll 2
        for (int e = tid; e < 64 * 64; e += kLrThreads) {
            const int r = e >> 6, c = e & 63;
            if (r < th && c < tw) {
                const int v = P.after[(size_t)(y0 + r) * P.after_stride + x0 + c];
                bad |= v > maxv;
                P.out[(size_t)(y0 + r) * P.out_stride + x0 + c] = (uint16_t)min(v, maxv);
            }
        }
        if (__syncthreads_or(bad) && tid == 0)
            raise_status(P.status, kFlagSampleRange);
        return;
    }

    // Stage Source::get (restoration.rs:233-244) over rows y0-3 .. y0+th+2,
    // columns x0-3 .. x0+tw+2.
    const int ss = rs.stripe_start, se = rs.stripe_end;
    const int srows = th + 6, scols = tw + 6;
    bool bad = false;
    if (tid < 256) {  // x / (x + 1) table of restoration.rs:330-336
        const int z = tid;
        xbx1[z] = z == 0 ? 1 : (uint16_t)(((z << 8) + z / 2) / (z + 1));
    }
#pragma unroll 2
    for (int e = tid; e < 70 * kS; e += kLrThreads) {
        const int r = e / kS, c = e - r * kS;
        if (r >= srows || c >= scols)
            continue;
        const int xc = min(max(x0 - 3 + c, 0), W - 1);
        const int yc = min(max(y0 - 3 + r, 0), H - 1);
        int v;
        if (yc < ss)
            v = P.before[(size_t)max(yc, ss - 2) * P.before_stride + xc];
        else if (yc > se)
            v = P.before[(size_t)min(yc, se + 2) * P.before_stride + xc];
        else
            v = P.after[(size_t)yc * P.after_stride + xc];
        bad |= v > maxv;
        S[r * kS + c] = (uint16_t)v;
    }
    if (__syncthreads_or(bad) && tid == 0)
        raise_status(P.status, kFlagSampleRange);

    if (u.kind == kUnitWiener) {
        // wiener, restoration.rs:264-283.
        const int round_h = bd == 12 ? 5 : 3;
        const int offset = 1 << (bd + 7 - round_h - 1);
        const int limit = (1 << (bd + 1 + 7 - round_h)) - 1;
        const int h0 = u.horizontal[0], h1 = u.horizontal[1], h2 = u.horizontal[2];
        const int h3 = 128 - 2 * (h0 + h1 + h2);
        const int v0 = u.vertical[0], v1 = u.vertical[1], v2 = u.vertical[2];
        const int v3 = 128 - 2 * (v0 + v1 + v2);
#pragma unroll 3
        for (int e = tid; e < srows * 64; e += kLrThreads) {
            const int r = e >> 6, c = e & 63;
            const uint16_t *s = &S[r * kS + c];
            int h = h0 * (s[0] + s[6]) + h1 * (s[1] + s[5]) + h2 * (s[2] + s[4]) + h3 * s[3];
            h = round2i(h, round_h);
            M[r * 64 + c] = min(max(h, -offset), limit - offset);
        }
        __syncthreads();
#pragma unroll 2
        for (int e = tid; e < 64 * 64; e += kLrThreads) {
            const int r = e >> 6, c = e & 63;
            if (r >= th || c >= tw)
                continue;
            const int *m = &M[r * 64 + c];
            int s = v0 * (m[0] + m[6 * 64]) + v1 * (m[64] + m[5 * 64]) +
                    v2 * (m[2 * 64] + m[4 * 64]) + v3 * m[3 * 64];
            s = round2i(s, 14 - round_h);
            P.out[(size_t)(y0 + r) * P.out_stride + x0 + c] = (uint16_t)min(max(s, 0), maxv);
        }
        return;
    }

    // self_guided, restoration.rs:373-384, with guided_pass 341-371.
    const int set = u.set & 15;
    const int r0 = kSgr[set][0], e0 = kSgr[set][1], r1 = kSgr[set][2], e1 = kSgr[set][3];
    uint16_t *VS = reinterpret_cast<uint16_t *>(M);
    int *VQ = M + 66 * kS / 2;
    int *A = AB, *B = AB + kG * kG;
    constexpr int kPx = 64 * 64 / kLrThreads;  // output samples per thread
    int f0[kPx];
    if (r0)
        sgr_grid<2>(S, VS, VQ, A, B, th, tw, y0, bd, e0, xbx1);
#pragma unroll
    for (int i = 0; i < kPx; i++) {
        const int e = tid + i * kLrThreads, r = e >> 6, c = e & 63;
        const int uu = S[(r + 3) * kS + c + 3];
        if (!r0 || r >= th || c >= tw) {
            f0[i] = uu << 4;
        } else if ((y0 + r) & 1) {  // odd row: its own row, weights 5 6 5, shift 4
            const int k = (r + 1) * kG + c;
            const int a = 5 * (A[k] + A[k + 2]) + 6 * A[k + 1];
            const int b = 5 * (B[k] + B[k + 2]) + 6 * B[k + 1];
            f0[i] = (a * uu + b + (1 << 7)) >> 8;
        } else {  // even row: the odd rows above and below, shift 5
            const int k = r * kG + c, k2 = k + 2 * kG;
            const int a = 5 * (A[k] + A[k + 2] + A[k2] + A[k2 + 2]) + 6 * (A[k + 1] + A[k2 + 1]);
            const int b = 5 * (B[k] + B[k + 2] + B[k2] + B[k2 + 2]) + 6 * (B[k + 1] + B[k2 + 1]);
            f0[i] = (a * uu + b + (1 << 8)) >> 9;
        }
    }
    if (r1) {
        if (r0)
            __syncthreads();  // pass-0 grids fully consumed
        sgr_grid<1>(S, VS, VQ, A, B, th, tw, y0, bd, e1, xbx1);
    }
    const int w0 = u.xqd[0], w1 = u.xqd[1];
#pragma unroll
    for (int i = 0; i < kPx; i++) {
        const int e = tid + i * kLrThreads, r = e >> 6, c = e & 63;
        if (r >= th || c >= tw)
            continue;
        const int uu = S[(r + 3) * kS + c + 3];
        int f1 = uu << 4;
        if (r1) {  // weights 4 on the cross, 3 on the diagonals, shift 5
            const int k = r * kG + c, k1 = k + kG, k2 = k + 2 * kG;
            const int a = 4 * (A[k + 1] + A[k1] + A[k1 + 1] + A[k1 + 2] + A[k2 + 1]) +
                          3 * (A[k] + A[k + 2] + A[k2] + A[k2 + 2]);
            const int b = 4 * (B[k + 1] + B[k1] + B[k1 + 1] + B[k1 + 2] + B[k2 + 1]) +
                          3 * (B[k] + B[k + 2] + B[k2] + B[k2 + 2]);
            f1 = (a * uu + b + (1 << 8)) >> 9;
        }
        int v = w0 * f0[i] + w1 * (uu << 4) + (128 - w0 - w1) * f1;
        v = (v + (1 << 10)) >> 11;
        P.out[(size_t)(y0 + r) * P.out_stride + x0 + c] = (uint16_t)min(max(v, 0), maxv);
    }
}

hipError_t launch_lr(const LrPlane *planes, int n, uint32_t blocks, hipStream_t s)
{
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_lr, dim3(blocks), dim3(kLrThreads), 0, s, planes, n);
    return hipGetLastError();
}

}  // namespace av1pf
