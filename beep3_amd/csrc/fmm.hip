// Far field by a Chebyshev-interpolation fast multipole method on a uniform
// sparse octree (gfx950).
//
// What it computes is exactly the quantity of the direct far-field kernel
// (kernels_far.hip): for every target patch i the 1-point centroid
// interaction with every other patch j, written as three potentials of
// monopole + dipole sources
//   Phi_L (Laplace 1/r):      monopole -a h,    dipole  ie a f n
//   Phi_1 (Yukawa e^-kr / r): monopole  a h,    dipole     -a f n
//   Phi_2 (Yukawa):           monopole ie a h,  dipole  -ie a f n
//   f_i += (Phi_L + Phi_1)(x_i) / 4pi,  h_i += n_i . grad(Phi_L + Phi_2)(x_i) / 4pi.
// This replaces the reference's 12-component Yukawa/Laplace plane-wave FMM
// (src/fmm/fmm_bem_hybrid.h:184-642): the same interaction, three expansions
// instead of twelve, every translation a small dense matrix.
//
// Expansions hold the values at the P^3 tensor Chebyshev nodes of each box:
//   P2M  M_m = sum_j [ q_j S_m(y_j) + d_j . grad S_m(y_j) ]      (leaf boxes)
//   M2M  M_parent = T_c M_child       (separable P x P factors, exact)
//   M2L  L_t += K(t) M_s              (K = kernel between node sets, one dense
//                                      NP x NP matrix per relative offset t;
//                                      Laplace scaled by 1/h, Yukawa per level)
//   L2L  L_child += T_c^T L_parent
//   L2P  Phi(x) = sum_n S_n(x) L_n,  grad Phi(x) = sum_n grad S_n(x) L_n
//   P2P  direct sum over the 27 neighbour leaf boxes (self excluded)
// with S_n(x) = prod_d S_P(xi_{n_d}, u_d), S_P(xi, u) = 1/P + 2/P sum_{k>=1}
// T_k(xi) T_k(u), u = box-local coordinate in [-1, 1].
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "device.hpp"
#include "fmm.hpp"
#include "pw.hpp"

namespace beepdev {

namespace {

constexpr double PI = 3.14159265358979323846;

// 1-D interpolation basis and derivative at u for all P nodes.
// tn[k * P + j] = T_j(xi_k).
template <typename T, int P>
__device__ __forceinline__ void basis1d(T u, const T* __restrict__ tn, T* s, T* ds) {
    T tj[P], dj[P];
    tj[0] = T(1);
    dj[0] = T(0);
    if (P > 1) {
        tj[1] = u;
        dj[1] = T(1);
    }
    // U_{j-1} recurrence for derivatives: T_j' = j U_{j-1}
    T um2 = T(1), um1 = T(2) * u;  // U_0, U_1
    for (int j = 2; j < P; ++j) {
        tj[j] = T(2) * u * tj[j - 1] - tj[j - 2];
        dj[j] = T(j) * um1;
        const T un = T(2) * u * um1 - um2;
        um2 = um1;
        um1 = un;
    }
    for (int k = 0; k < P; ++k) {
        T a = T(0), b = T(0);
        for (int j = 1; j < P; ++j) {
            a += tn[k * P + j] * tj[j];
            b += tn[k * P + j] * dj[j];
        }
        s[k] = T(1) / T(P) + T(2) / T(P) * a;
        ds[k] = T(2) / T(P) * b;
    }
}

template <typename T>
struct alignas(16) SrcR {
    T x, y, z, ie;
    T dx, dy, dz, sh;
};

// ---------------------------------------------------------------------------
// P2M over leaf boxes.  One block per leaf box; thread n = node.
template <typename T, int P>
__global__ __launch_bounds__(256) void k_p2m(const FmmDev d, const Src64* __restrict__ src,
                                             T* __restrict__ M) {
    constexpr int NP = P * P * P;
    constexpr int CH = 64;
    __shared__ T sb[CH][3][P], dsb[CH][3][P];
    __shared__ T q[CH][4];  // sh, ie, (dv . grad handled through dvx..)
    __shared__ T dvv[CH][3];
    __shared__ T tn[P * P];
    const int bl = blockIdx.x;                  // leaf box (local index at leaf level)
    const int box = d.leaf_first + bl;
    const int s0 = d.leaf_src[bl], s1 = d.leaf_src[bl + 1];
    const double cx = d.box_c[box].x, cy = d.box_c[box].y, cz = d.box_c[box].z;
    const double inv_half = 2.0 / d.box_h[d.D];
    const int t = threadIdx.x;
    for (int k = t; k < P * P; k += blockDim.x) tn[k] = T(d.tn[k]);
    const int a = t / (P * P), b = (t / P) % P, c = t % P;
    T mL = 0, m1 = 0, m2 = 0;
    const T g2 = T(inv_half);
    for (int j0 = s0; j0 < s1; j0 += CH) {
        __syncthreads();
        if (t < CH && j0 + t < s1) {
            const Src64 s = src[j0 + t];
            T sx[P], dx[P];
            basis1d<T, P>(T((s.x - cx) * inv_half), tn, sx, dx);
            for (int k = 0; k < P; ++k) { sb[t][0][k] = sx[k]; dsb[t][0][k] = dx[k]; }
            basis1d<T, P>(T((s.y - cy) * inv_half), tn, sx, dx);
            for (int k = 0; k < P; ++k) { sb[t][1][k] = sx[k]; dsb[t][1][k] = dx[k]; }
            basis1d<T, P>(T((s.z - cz) * inv_half), tn, sx, dx);
            for (int k = 0; k < P; ++k) { sb[t][2][k] = sx[k]; dsb[t][2][k] = dx[k]; }
            q[t][0] = T(s.sh);
            q[t][1] = T(s.inv_eps);
            dvv[t][0] = T(s.dx);
            dvv[t][1] = T(s.dy);
            dvv[t][2] = T(s.dz);
        }
        __syncthreads();
        const int m = min(CH, s1 - j0);
        if (t < NP) {
            for (int j = 0; j < m; ++j) {
                const T sx = sb[j][0][a], sy = sb[j][1][b], sz = sb[j][2][c];
                const T syz = sy * sz;
                const T s = sx * syz;
                const T gx = dsb[j][0][a] * syz;
                const T gy = sx * dsb[j][1][b] * sz;
                const T gz = sx * sy * dsb[j][2][c];
                const T dg = g2 * (dvv[j][0] * gx + dvv[j][1] * gy + dvv[j][2] * gz);
                const T sh = q[j][0], ie = q[j][1];
                mL += ie * dg - sh * s;
                const T c1 = sh * s - dg;
                m1 += c1;
                m2 += ie * c1;
            }
        }
    }
    if (t < NP) {
        T* out = M + size_t(box) * 3 * NP;
        out[t] = mL;
        out[NP + t] = m1;
        out[2 * NP + t] = m2;
    }
}

// M2M: one block per parent box (boxes [first, first+count)).  The child
// -> parent interpolation is a tensor product of one P x P factor per axis,
// applied one axis at a time (3 P^4 instead of P^6 products per potential),
// all children side by side; the children are then added in list order.
template <typename T, int P>
__global__ __launch_bounds__(256) void k_m2m(const FmmDev d, int first, T* __restrict__ M) {
    constexpr int NP = P * P * P, NV = 3 * NP;
    __shared__ T tt[2][P][P];
    __shared__ T buf[2][8][NV];
    __shared__ int soct[8];
    const int box = first + blockIdx.x;
    const int t = threadIdx.x;
    const int k0 = d.child_ptr[box], nch = d.child_ptr[box + 1] - k0;  // at most 8
    for (int k = t; k < 2 * P * P; k += blockDim.x) (&tt[0][0][0])[k] = T(d.t_oct[k]);
    if (t < nch) soct[t] = d.child_oct[k0 + t];
    for (int q = t; q < nch * NV; q += blockDim.x) {
        const int k = q / NV, r = q - k * NV;
        buf[0][k][r] = M[size_t(d.child_box[k0 + k]) * NV + r];
    }
    __syncthreads();
    // z: buf1[k][e][a2][b2][c] = sum_c2 tt[oz][c][c2] buf0[k][e][a2][b2][c2]
    for (int q = t; q < nch * NV; q += blockDim.x) {
        const int k = q / NV, r = q - k * NV;
        const int c = r % P, base = r - c;
        const int oz = (soct[k] >> 2) & 1;
        T v = 0;
#pragma unroll
        for (int c2 = 0; c2 < P; ++c2) v += tt[oz][c][c2] * buf[0][k][base + c2];
        buf[1][k][r] = v;
    }
    __syncthreads();
    // y: buf0[k][e][a2][b][c] = sum_b2 tt[oy][b][b2] buf1[k][e][a2][b2][c]
    for (int q = t; q < nch * NV; q += blockDim.x) {
        const int k = q / NV, r = q - k * NV;
        const int b = (r / P) % P, base = r - b * P;
        const int oy = (soct[k] >> 1) & 1;
        T v = 0;
#pragma unroll
        for (int b2 = 0; b2 < P; ++b2) v += tt[oy][b][b2] * buf[1][k][base + b2 * P];
        buf[0][k][r] = v;
    }
    __syncthreads();
    // x, and the sum over children: M[e][a][b][c] = sum_k sum_a2 tt[ox][a][a2] buf0[k][e][a2][b][c]
    for (int r = t; r < NV; r += blockDim.x) {
        const int a = (r / (P * P)) % P, base = r - a * P * P;
        T v = 0;
        for (int k = 0; k < nch; ++k) {
            const int ox = soct[k] & 1;
#pragma unroll
            for (int a2 = 0; a2 < P; ++a2) v += tt[ox][a][a2] * buf[0][k][base + a2 * P * P];
        }
        M[size_t(box) * NV + r] = v;
    }
}

// M2L on the matrix cores.  The kernel between the node sets of a source
// box at offset t (box units, max |t_k| >= 2) from its target depends on t
// only through |t + (xi_m - xi_n)/2|; the Chebyshev node set is symmetric
// (xi_{P-1-a} = -xi_a), so for t = sigma t_c with sigma a signed permutation
// of the axes and t_c one of 16 canonical offsets (t_c0 >= t_c1 >= t_c2 >= 0)
//   K_t[m][n] = K_c[phi(m)][phi(n)],
// phi = the same signed permutation acting on node indices.  The plan groups
// the pairs of a level by canonical offset into chunks of 64; a block stages
// the source expansions of its chunk through LDS with each pair's node
// permutation applied (B operands), multiplies by the one canonical table
// (A operands, 16 tables per level: they stay in L2) with 16x16x4 MFMA tiles.
// Pairs of a class are listed target by target, so a chunk holds runs of one
// target's pairs (segments, 3.3 pairs on average): the block sums each
// segment's results, permuted back to the target's node order, through LDS
// and writes one slot per segment; k_m2l_sum adds a target's segment slots
// (deterministic, no atomics).
typedef double mf64x4 __attribute__((ext_vector_type(4)));
typedef float mf32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    using acc_t = mf64x4;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator element r of lane l holds row l/16 + 4r (column l % 16)
    static __device__ __forceinline__ int row(int l, int r) { return l / 16 + 4 * r; }
};
template <>
struct Mfma<float> {
    using acc_t = mf32x4;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator element r of lane l holds row 4 (l/16) + r (column l % 16)
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l / 16) + r; }
};

// actual node index of canonical node index mc for a pair with symmetry code
// sym: bits 0-5 the actual axis of canonical axes 0, 1, 2 (2 bits each),
// bits 6-8 a reflection of canonical axis 0, 1, 2
template <int P>
inline int node_of(int mc, int sym) {
    const int c[3] = {mc / (P * P), (mc / P) % P, mc % P};
    int a[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int ax = (sym >> (2 * k)) & 3;
        const int v = ((sym >> (6 + k)) & 1) ? P - 1 - c[k] : c[k];
        a[ax] = v;
    }
    return (a[0] * P + a[1]) * P + a[2];
}

constexpr int M2L_G = 64;     // pairs per block (4 MFMA column tiles)
constexpr int M2L_WAVES = 8;  // waves per block, each owning output row tiles
// K-dimension slab staged in LDS per barrier: 32 (FP64) or 64 (FP32)
// K-indices, measured fastest; rows padded by 4 (spreads the B reads over the
// banks, keeps 16-byte alignment)
template <typename T>
constexpr int m2l_slab() { return sizeof(T) == 8 ? 32 : 64; }

template <typename T, int P>
__global__ __launch_bounds__(512) void k_m2l(const FmmDev d, const int4* __restrict__ chunks,
                                             const T* __restrict__ kl,
                                             const T* const* __restrict__ ytabs,
                                             const T* __restrict__ M, T* __restrict__ Cout) {
    constexpr int NP = P * P * P;
    constexpr int NPP = (NP + 15) / 16 * 16;
    constexpr int NTILE = NPP / 16;
    constexpr int NTW = (NTILE + M2L_WAVES - 1) / M2L_WAVES;  // row tiles per wave
    constexpr int G = M2L_G, CT = G / 16, MS = m2l_slab<T>(), MSP = MS + 4;
    constexpr int NQ = (G * MS + 511) / 512;  // slab entries per thread
    constexpr int NSLAB = (NPP + MS - 1) / MS;
    constexpr int NSTEP = NPP / 4, SPS = MS / 4;  // k-steps, k-steps per slab
    constexpr int RV = 16 / int(sizeof(T));       // k-steps per 16-byte LDS read
    static_assert(SPS % RV == 0 && NSTEP % RV == 0, "slab / vector shape");
    using acc_t = typename Mfma<T>::acc_t;
    typedef T vec_t __attribute__((ext_vector_type(RV)));
    // two B slab buffers [g][position of m] (slab s + 1 is stored while slab
    // s feeds the matrix cores), then the results Y[g][canonical node] of
    // the block's pairs, rows padded by one (conflict-free column writes)
    constexpr int SLAB_BUF = G * MSP, YS = NPP + 1;
    constexpr int SMEM = 2 * SLAB_BUF > G * YS ? 2 * SLAB_BUF : G * YS;
    __shared__ alignas(16) T smem[SMEM];
    __shared__ int psrc[G];
    __shared__ const unsigned short* pperm[G];
    __shared__ const unsigned short* pphi[G];
    __shared__ int sstart[G + 1];  // first pair of each segment of the chunk (+ end)
    // block = (chunk, potential): potential 0 uses the Laplace table,
    // potentials 1 and 2 the Yukawa table of the level
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int cid = bid / 3, e = bid - 3 * cid;
    const int4 ch = chunks[cid];  // canonical offset, pairs, level
    const int off = ch.x, p0 = ch.y, np = ch.z - ch.y, level = ch.w;
    const int t = threadIdx.x;
    const int seg0 = d.chunk_seg[cid], nseg = d.chunk_seg[cid + 1] - seg0;
    if (t < G) {
        const size_t sym = size_t(t < np ? d.pair_sym[p0 + t] : 0) * NP;
        psrc[t] = t < np ? d.pair_src[p0 + t] : 0;
        pperm[t] = d.psi + sym;
        pphi[t] = d.phi + sym;
        if (t < np) {
            const int sg = d.pair_seg[p0 + t];
            if (t == 0 || d.pair_seg[p0 + t - 1] != sg) sstart[sg - seg0] = t;
        }
        if (t == 0) sstart[nseg] = np;
    }
    const int w = t >> 6, l = t & 63;
    const int kl16 = l / 16, il = l % 16;
    const T* KT = (e == 0 ? kl : ytabs[level]) + size_t(off) * NP * NP;
    acc_t acc[NTW][CT];
#pragma unroll
    for (int k = 0; k < NTW; ++k)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[k][c] = acc_t{0, 0, 0, 0};
    // B slabs: B[g][mm] = M_src(g)[e][psi_g(m0 + mm)], the source expansions
    // in the canonical node order of their pairs.  Thread t owns the same
    // (g, mm) entries of every slab (consecutive threads walk mm of one g, a
    // run inside one expansion).  In each g row, m = 4 s + k (k-step s of the
    // slab, K index k of the MFMA operand) sits at k * SPS + s, so a lane
    // finds the operands of RV consecutive k-steps side by side.
    __syncthreads();  // pair lists written
    int qsrc[NQ], qmm[NQ];
    const unsigned short* qperm[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int q = t + 512 * i;
        const int g = q / MS, mm = q - g * MS;
        const bool ok = q < G * MS && g < np;
        qsrc[i] = ok ? psrc[g] * 3 * NP + e * NP : -1;
        qperm[i] = ok ? pperm[g] : nullptr;
        qmm[i] = q < G * MS ? g * MSP + (mm & 3) * SPS + (mm >> 2) : -1;
    }
    auto fetch = [&](int m0, T* v) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int m = m0 + (t + 512 * i) % MS;
            v[i] = (qsrc[i] >= 0 && m < NP) ? M[size_t(qsrc[i]) + qperm[i][m]] : T(0);
        }
    };
    T nxt[NQ];
    fetch(0, nxt);
#pragma unroll
    for (int i = 0; i < NQ; ++i)
        if (qmm[i] >= 0) smem[qmm[i]] = nxt[i];
    if (NSLAB > 1) fetch(MS, nxt);  // in flight during the first slab
    // A operands (kernel table values) one k-step ahead: the loads of step
    // k+1 are in flight while the MFMAs of step k run
    T xa_n[NTW];
    auto load_a = [&](int step, T* xa) {
        const int m = 4 * step + kl16;  // K-dimension index (canonical source node)
#pragma unroll
        for (int k = 0; k < NTW; ++k) {
            const int n = (w + M2L_WAVES * k) * 16 + il;  // output row (canonical target node)
            xa[k] = (m < NP && n < NP) ? KT[m * NP + n] : T(0);
        }
    };
    load_a(0, xa_n);
#pragma unroll 1
    for (int sg = 0; sg < NSTEP; sg += RV) {
        const int ss = sg % SPS, slab = sg / SPS;  // first k-step of the group in its slab
        if (ss == 0) {
            // slab `slab` is complete in its buffer, and every read of the
            // previous slab (the other buffer) is done
            __syncthreads();
            if (slab + 1 < NSLAB) {
#pragma unroll
                for (int i = 0; i < NQ; ++i)
                    if (qmm[i] >= 0) smem[((slab + 1) & 1) * SLAB_BUF + qmm[i]] = nxt[i];
                if (slab + 2 < NSLAB) fetch((slab + 2) * MS, nxt);
            }
        }
        const T* sbuf = smem + (slab & 1) * SLAB_BUF;
        vec_t bv[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c)
            bv[c] = *reinterpret_cast<const vec_t*>(&sbuf[(c * 16 + il) * MSP + kl16 * SPS + ss]);
#pragma unroll
        for (int r = 0; r < RV; ++r) {
            const int step = sg + r;
            T xa[NTW];
#pragma unroll
            for (int k = 0; k < NTW; ++k) xa[k] = xa_n[k];
            if (step + 1 < NSTEP) load_a(step + 1, xa_n);
#pragma unroll
            for (int k = 0; k < NTW; ++k) {
                if (w + M2L_WAVES * k >= NTILE) continue;
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[k][c] = Mfma<T>::op(xa[k], bv[c][r], acc[k][c]);
            }
        }
    }
    // the results through LDS: Y[g][n] of pair g at canonical target node n
    // (accumulator element r of lane l holds row Mfma<T>::row(l, r) of the
    // tile, column l % 16)
    const T scale = e == 0 ? T(1.0 / d.box_h[level]) : T(1);  // Laplace tables are built at h = 1
    __syncthreads();  // every read of the last slab done: the slab buffers become Y
#pragma unroll
    for (int k = 0; k < NTW; ++k) {
        const int nt = w + M2L_WAVES * k;
        if (nt >= NTILE) continue;
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const acc_t v = acc[k][c];
            T* col = smem + (c * 16 + il) * YS + nt * 16;
#pragma unroll
            for (int r = 0; r < 4; ++r) col[Mfma<T>::row(l, r)] = v[r] * scale;
        }
    }
    __syncthreads();
    // one slot per segment: its pairs' results at each actual target node m
    // (canonical node phi_g(m) of pair g), added in pair order
    for (int o = t; o < nseg * NP; o += 512) {
        const int sl = o / NP, m = o - sl * NP;
        T sum = 0;
        for (int g = sstart[sl]; g < sstart[sl + 1]; ++g) sum += smem[g * YS + pphi[g][m]];
        Cout[(size_t(seg0 + sl) * 3 + e) * NP + m] = sum;
    }
}

// Sum the M2L segment slots of every target box into its local expansion
// (one block per box at levels >= 2; a box's segments listed in increasing
// order in tgt_ptr / tgt_seg).  Each thread owns outputs (e, n) and adds the
// slots in list order, so the sum is deterministic.
template <typename T, int P>
__global__ __launch_bounds__(256) void k_m2l_sum(const FmmDev d, int first_box,
                                                 const T* __restrict__ Cin, T* __restrict__ L) {
    constexpr int NP = P * P * P;
    constexpr int NQ = (3 * NP + 255) / 256;
    const int box = first_box + blockIdx.x;
    const int k0 = d.tgt_ptr[box], k1 = d.tgt_ptr[box + 1];
    if (k0 == k1) return;
    const int t = threadIdx.x;
    T acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = T(0);
#pragma unroll 4
    for (int k = k0; k < k1; ++k) {
        const T* c = Cin + size_t(d.tgt_seg[k]) * 3 * NP;
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            if (t + 256 * i < 3 * NP) acc[i] += c[t + 256 * i];
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i)
        if (t + 256 * i < 3 * NP) L[size_t(box) * 3 * NP + t + 256 * i] += acc[i];
}

// L2L: one block per child box at level l+1 (boxes [first, first+count)).
// The parent -> child interpolation, one axis at a time as in k_m2m.
template <typename T, int P>
__global__ __launch_bounds__(256) void k_l2l(const FmmDev d, int first, T* __restrict__ L) {
    constexpr int NP = P * P * P, NV = 3 * NP;
    __shared__ T buf[2][NV];
    __shared__ T tt[2][P][P];
    const int box = first + blockIdx.x;
    const int par = d.box_parent[box];
    const int oct = d.box_oct[box];
    const int ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1;
    const int t = threadIdx.x;
    for (int k = t; k < 2 * P * P; k += blockDim.x) (&tt[0][0][0])[k] = T(d.t_oct[k]);
    for (int q = t; q < NV; q += blockDim.x) buf[0][q] = L[size_t(par) * NV + q];
    __syncthreads();
    // z: buf1[e][a][b][c2] = sum_c tt[oz][c][c2] buf0[e][a][b][c]
    for (int r = t; r < NV; r += blockDim.x) {
        const int c2 = r % P, base = r - c2;
        T v = 0;
#pragma unroll
        for (int c = 0; c < P; ++c) v += tt[oz][c][c2] * buf[0][base + c];
        buf[1][r] = v;
    }
    __syncthreads();
    // y: buf0[e][a][b2][c2] = sum_b tt[oy][b][b2] buf1[e][a][b][c2]
    for (int r = t; r < NV; r += blockDim.x) {
        const int b2 = (r / P) % P, base = r - b2 * P;
        T v = 0;
#pragma unroll
        for (int b = 0; b < P; ++b) v += tt[oy][b][b2] * buf[1][base + b * P];
        buf[0][r] = v;
    }
    __syncthreads();
    // x: L_child[e][a2][b2][c2] += sum_a tt[ox][a][a2] buf0[e][a][b2][c2]
    for (int r = t; r < NV; r += blockDim.x) {
        const int a2 = (r / (P * P)) % P, base = r - a2 * P * P;
        T v = 0;
#pragma unroll
        for (int a = 0; a < P; ++a) v += tt[ox][a][a2] * buf[0][base + a * P * P];
        L[size_t(box) * NV + r] += v;
    }
}

// L2P: potentials and gradients of the local expansions at the targets of
// one leaf box.  One wave per leaf box; the box's expansions and the
// Chebyshev table sit in LDS, lane = target.
template <typename T, int P>
__global__ __launch_bounds__(64) void k_l2p(const FmmDev d, const Patch64* __restrict__ patches,
                                            const T* __restrict__ L, double* __restrict__ y,
                                            int n) {
    constexpr int NP = P * P * P;
    __shared__ T lb[3 * NP];
    __shared__ T tn[P * P];
    const int bl = xcd_remap(blockIdx.x, gridDim.x);
    const int box = d.leaf_first + bl;
    const int s0 = d.leaf_src[bl], s1 = d.leaf_src[bl + 1];
    const double cx = d.box_c[box].x, cy = d.box_c[box].y, cz = d.box_c[box].z;
    const T ih = T(2.0 / d.box_h[d.D]);
    const int t = threadIdx.x;
    for (int k = t; k < P * P; k += 64) tn[k] = T(d.tn[k]);
    for (int q = t; q < 3 * NP; q += 64) lb[q] = L[size_t(box) * 3 * NP + q];
    __syncthreads();
    for (int ti = s0 + t; ti < s1; ti += 64) {
        const int i = d.perm[ti];
        const Patch64 p = patches[i];
        T sx[P], dsx[P], sy[P], dsy[P], sz[P], dsz[P];
        basis1d<T, P>(T(p.cx - cx) * ih, tn, sx, dsx);
        basis1d<T, P>(T(p.cy - cy) * ih, tn, sy, dsy);
        basis1d<T, P>(T(p.cz - cz) * ih, tn, sz, dsz);
        T pL = 0, p1 = 0, gLx = 0, gLy = 0, gLz = 0, g2x = 0, g2y = 0, g2z = 0;
#pragma unroll 1
        for (int a = 0; a < P; ++a)
            for (int b = 0; b < P; ++b) {
                const T sab = sx[a] * sy[b], dab = dsx[a] * sy[b], adb = sx[a] * dsy[b];
                for (int c = 0; c < P; ++c) {
                    const int m = (a * P + b) * P + c;
                    const T s = sab * sz[c];
                    const T gx = dab * sz[c], gy = adb * sz[c], gz = sab * dsz[c];
                    const T l0 = lb[m], l1 = lb[NP + m], l2 = lb[2 * NP + m];
                    pL += s * l0;
                    p1 += s * l1;
                    gLx += gx * l0; gLy += gy * l0; gLz += gz * l0;
                    g2x += gx * l2; g2y += gy * l2; g2z += gz * l2;
                }
            }
        const T nx = T(p.nx), ny = T(p.ny), nz = T(p.nz);
        y[i] += double(pL + p1) * FOUR_PI_INV;
        y[n + i] += double(ih * (nx * (gLx + g2x) + ny * (gLy + g2y) + nz * (gLz + g2z))) *
                    FOUR_PI_INV;
    }
}

template <typename T>
struct P2PCap {
    static constexpr int value = 256;  // 16 KB of FP64 sources
};
template <>
struct P2PCap<float> {
    static constexpr int value = 512;  // 16 KB of FP32 sources
};

// P2P: direct sum over the sources of the (at most 27) neighbour leaf boxes,
// self excluded.  One block per leaf box, 256 threads = 16 groups of 16
// lanes.  The neighbour sources are staged in LDS in passes of CAP (positions
// relative to the box centre, formed in FP64); a group takes a chunk of 16
// targets (lane = target) and every ns-th staged source, so small leaves
// still keep all lanes busy; the ns partial sums of a chunk are combined in
// LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_p2p(const FmmDev d, const Src64* __restrict__ src,
                                             const Patch64* __restrict__ patches, T kappa,
                                             double* __restrict__ y, int n) {
    constexpr int CAP = P2PCap<T>::value;
    __shared__ SrcR<T> tile[CAP];
    __shared__ int pre[28];
    __shared__ int beg[27];
    __shared__ T red[2][16][16];
    const int bl = xcd_remap(blockIdx.x, gridDim.x);
    const int box = d.leaf_first + bl;
    const int s0 = d.leaf_src[bl], s1 = d.leaf_src[bl + 1];
    const double cx = d.box_c[box].x, cy = d.box_c[box].y, cz = d.box_c[box].z;
    const int t = threadIdx.x;
    const int k0 = d.nb_ptr[bl], nk = d.nb_ptr[bl + 1] - k0;  // <= 27 by construction
    if (t == 0) {
        int acc = 0;
        for (int k = 0; k < nk; ++k) {
            const int nbl = d.nb_leaf[k0 + k];
            beg[k] = d.leaf_src[nbl];
            pre[k] = acc;
            acc += d.leaf_src[nbl + 1] - d.leaf_src[nbl];
        }
        pre[nk] = acc;
    }
    __syncthreads();
    const int S = pre[nk];
    const int g = t >> 4, lane = t & 15;
    const int nch = (s1 - s0 + 15) >> 4;
    // nch < 16: group g works on chunk g % nch with source phase g / nch of ns;
    // nch >= 16: one phase, group g takes chunks g, g + 16, ... (one per round)
    const int ns = nch < 16 ? 16 / nch : 1;
    const int rounds = nch < 16 ? 1 : (nch + 15) >> 4;
    const bool active = g < nch * ns || nch >= 16;
    const int phase = nch < 16 ? g / nch : 0;
    for (int r = 0; r < rounds; ++r) {
        const int c = nch < 16 ? g % nch : g + 16 * r;
        const int ti = s0 + 16 * c + lane;
        const bool valid = active && c < nch && ti < s1;
        T xi = 0, yi = 0, zi = 0, n0x = 0, n0y = 0, n0z = 0;
        if (valid) {
            const Patch64 p = patches[d.perm[ti]];
            xi = T(p.cx - cx); yi = T(p.cy - cy); zi = T(p.cz - cz);
            n0x = T(p.nx); n0y = T(p.ny); n0z = T(p.nz);
        }
        T af = 0, ah = 0;
        for (int base = 0; base < S; base += CAP) {
            const int m = min(CAP, S - base);
            __syncthreads();
            for (int q = t; q < m; q += 256) {
                const int pidx = base + q;
                int k = 0;
                while (pre[k + 1] <= pidx) ++k;
                const Src64 sv = src[beg[k] + (pidx - pre[k])];
                tile[q] = SrcR<T>{T(sv.x - cx), T(sv.y - cy), T(sv.z - cz), T(sv.inv_eps),
                                  T(sv.dx), T(sv.dy), T(sv.dz), T(sv.sh)};
            }
            __syncthreads();
            if (valid) {
                for (int jj = phase; jj < m; jj += ns) {
                    const SrcR<T> q = tile[jj];
                    p2p_pair<T>(q.x - xi, q.y - yi, q.z - zi, q.ie, q.dx, q.dy, q.dz, q.sh, n0x,
                                n0y, n0z, kappa, af, ah);
                }
            }
        }
        if (ns > 1) {
            red[0][g][lane] = af;
            red[1][g][lane] = ah;
            __syncthreads();
            if (valid && phase == 0) {
                for (int k = 1; k < ns; ++k) {
                    af += red[0][g + k * nch][lane];
                    ah += red[1][g + k * nch][lane];
                }
            }
        }
        if (valid && phase == 0) {
            const int i = d.perm[ti];
            y[i] += double(af) * FOUR_PI_INV;
            y[n + i] += double(ah) * FOUR_PI_INV;
        }
    }
}

template <typename T>
__global__ void k_zero(size_t n, T* __restrict__ p) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = T(0);
}

// Kernel tables between node sets: out[c][m][n] = G(|t_c + (xi_m - xi_n)/2| * h)
// (m = source node, n = target node) for the canonical offsets t_c (box
// units, t_c0 >= t_c1 >= t_c2 >= 0, t_c0 in {2, 3}); laplace: 1/r with h = 1
// (scaled by 1/h at use), else e^{-kappa r}/r at box size h.
template <typename T, int P>
__global__ void k_ktab(const double* __restrict__ xi, const int4* __restrict__ canon, int ncanon,
                       double h, double kappa, int laplace, T* __restrict__ out) {
    constexpr int NP = P * P * P;
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= size_t(ncanon) * NP * NP) return;
    const int off = int(idx / (size_t(NP) * NP));
    const int rem = int(idx - size_t(off) * NP * NP);
    const int m = rem / NP, nn = rem - m * NP;
    const int tx = canon[off].x, ty = canon[off].y, tz = canon[off].z;
    const int ma = m / (P * P), mb = (m / P) % P, mc = m % P;
    const int na = nn / (P * P), nb = (nn / P) % P, nc = nn % P;
    const double vx = tx + 0.5 * (xi[ma] - xi[na]);
    const double vy = ty + 0.5 * (xi[mb] - xi[nb]);
    const double vz = tz + 0.5 * (xi[mc] - xi[nc]);
    const double r = sqrt(vx * vx + vy * vy + vz * vz);
    double v = 0;
    if (max(abs(tx), max(abs(ty), abs(tz))) > 1) {
        if (laplace) v = 1.0 / r;
        else v = exp(-kappa * h * r) / (h * r);
    }
    out[idx] = T(v);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side: plan construction and launches.

namespace {

uint64_t bkey(uint32_t x, uint32_t y, uint32_t z) {
    return (uint64_t(x) << 42) | (uint64_t(y) << 21) | uint64_t(z);
}

// The 16 canonical M2L offsets (a, b, c), 3 >= a >= b >= c >= 0, a >= 2.
const std::vector<std::array<int, 3>>& canonical_list() {
    static const std::vector<std::array<int, 3>> list = [] {
        std::vector<std::array<int, 3>> v;
        for (int a = 2; a <= 3; ++a)
            for (int b = 0; b <= a; ++b)
                for (int c = 0; c <= b; ++c) v.push_back({a, b, c});
        return v;
    }();
    return list;
}

// Canonical index of offset t and the symmetry code of the node map (see
// node_of): canonical axis k is actual axis ax[k] with |t| sorted
// descending, reflected where that component of t is negative.
int canonical_offset(int tx, int ty, int tz, int* sym);

// The 48 signed axis permutations as node-map codes (see node_of), in a
// fixed order; a pair stores the index of its code.
const std::vector<int>& symmetry_codes() {
    static const std::vector<int> codes = [] {
        std::vector<int> v;
        const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (const auto& pm : perms)
            for (int sgn = 0; sgn < 8; ++sgn) v.push_back(pm[0] | (pm[1] << 2) | (pm[2] << 4) | (sgn << 6));
        return v;
    }();
    return codes;
}

int symmetry_index(int code) {
    const auto& c = symmetry_codes();
    for (size_t k = 0; k < c.size(); ++k)
        if (c[k] == code) return int(k);
    throw std::runtime_error("bad M2L symmetry code");
}

int canonical_offset(int tx, int ty, int tz, int* sym) {
    const int t[3] = {tx, ty, tz};
    int ax[3] = {0, 1, 2};
    std::stable_sort(ax, ax + 3, [&](int p, int q) { return std::abs(t[p]) > std::abs(t[q]); });
    const std::array<int, 3> c = {std::abs(t[ax[0]]), std::abs(t[ax[1]]), std::abs(t[ax[2]])};
    int code = ax[0] | (ax[1] << 2) | (ax[2] << 4);
    for (int k = 0; k < 3; ++k)
        if (t[ax[k]] < 0) code |= 1 << (6 + k);
    *sym = code;
    const auto& L = canonical_list();
    for (size_t k = 0; k < L.size(); ++k)
        if (L[k] == c) return int(k);
    throw std::runtime_error("M2L offset outside the interaction range");
}

// canonical offset and symmetry index of every M2L offset t in [-3, 3]^3
// (max |t_k| >= 2), looked up per pair while the plan is built
struct OffsetClass {
    int canon = -1, sym = -1;
};
const std::array<OffsetClass, 343>& offset_classes() {
    static const std::array<OffsetClass, 343> tab = [] {
        std::array<OffsetClass, 343> a{};
        for (int tx = -3; tx <= 3; ++tx)
            for (int ty = -3; ty <= 3; ++ty)
                for (int tz = -3; tz <= 3; ++tz) {
                    if (std::max(std::abs(tx), std::max(std::abs(ty), std::abs(tz))) < 2) continue;
                    int code = 0;
                    const int c = canonical_offset(tx, ty, tz, &code);
                    a[size_t(((tx + 3) * 7 + (ty + 3)) * 7 + (tz + 3))] = {c, symmetry_index(code)};
                }
        return a;
    }();
    return tab;
}

// Boxes of one level: cell -> box number, a dense table when the level has
// at most 2^21 cells, else a hash map
struct BoxIndex {
    int lev = 0;
    std::vector<int> dense;
    std::unordered_map<uint64_t, int> sparse;
    explicit BoxIndex(int l = 0) : lev(l) {
        if (3 * l <= 21) dense.assign(size_t(1) << (3 * l), -1);
    }
    size_t cell(uint32_t x, uint32_t y, uint32_t z) const {
        return (size_t(x) << (2 * lev)) | (size_t(y) << lev) | size_t(z);
    }
    int find(uint32_t x, uint32_t y, uint32_t z) const {
        if (!dense.empty()) return dense[cell(x, y, z)];
        const auto it = sparse.find(bkey(x, y, z));
        return it == sparse.end() ? -1 : it->second;
    }
    // true if the cell had no box yet (it now has v)
    bool insert(uint32_t x, uint32_t y, uint32_t z, int v) {
        if (!dense.empty()) {
            int& r = dense[cell(x, y, z)];
            if (r >= 0) return false;
            r = v;
            return true;
        }
        return sparse.emplace(bkey(x, y, z), v).second;
    }
};

template <typename T>
void upload(DevArena& ar, const char* name, T*& dst, const std::vector<T>& v, hipStream_t s) {
    dst = ar.get<T>(name, std::max<size_t>(v.size(), 1));
    if (!v.empty())
        hip_check(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s),
                  "H2D fmm");
}

}  // namespace

// Host data carried from fmm_plan_begin to fmm_plan_finish.
struct FmmPlanStage {
    double x0[3] = {0, 0, 0};
    std::vector<uint32_t> cx, cy, cz;
    std::vector<unsigned> ck;  // packed cells, until the upload has completed
};

FmmPlan* fmm_plan_begin(int N, const double* centroid, int order, int leaf_target,
                        DevArena& arena, hipStream_t s, int engine) {
    if (engine == 0 && (order < 3 || order > 6)) throw std::runtime_error("FMM order must be 3..6");
    if (engine == 1 && !pw_order_supported(order))
        throw std::runtime_error("plane-wave FMM order must be 9, 12 or 18");
    if (engine != 0 && engine != 1) throw std::runtime_error("unknown FMM engine");
    auto* pl = new FmmPlan();
    pl->arena = &arena;
    pl->N = N;
    pl->engine = engine;
    // the Chebyshev order; the plane-wave engine keeps its own
    pl->P = engine == 0 ? order : 3;
    if (engine == 1) {
        pl->pw_p = order;
        pl->pw_nlam = order <= 9 ? 9 : 18;
        pl->pw_jp = ((order + 1) * (order + 1) + 15) / 16 * 16;
    }
    pl->stage = new FmmPlanStage();
    // bounding cube
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], centroid[3 * i + k]);
            hi[k] = std::max(hi[k], centroid[3 * i + k]);
        }
    double edge = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    edge = edge * (1.0 + 1e-9) + 1e-9;
    const double ctr[3] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
    double x0[3];
    for (int k = 0; k < 3; ++k) x0[k] = ctr[k] - edge / 2;
    // leaf level: first level whose mean occupancy of non-empty boxes <= target
    auto cell = [&](int i, int lev, int k) {
        const long cells = 1L << lev;
        long v = long(std::floor((centroid[3 * i + k] - x0[k]) / edge * cells));
        return uint32_t(std::min(cells - 1, std::max(0L, v)));
    };
    int D = 0;
    for (; D < 12; ++D) {
        // occupied boxes: a bitmap over the level's cells up to 2^24 of them
        // (N inserts into a hash map per level cost tens of ms at 10^5 patches)
        size_t occupied = 0;
        if (D <= 8) {
            const size_t cells = size_t(1) << (3 * D);
            std::vector<uint64_t> bits((cells + 63) / 64, 0);
            for (int i = 0; i < N; ++i) {
                const size_t c = (size_t(cell(i, D, 0)) << (2 * D)) | (size_t(cell(i, D, 1)) << D) |
                                 size_t(cell(i, D, 2));
                const uint64_t m = uint64_t(1) << (c & 63);
                if (!(bits[c >> 6] & m)) {
                    bits[c >> 6] |= m;
                    ++occupied;
                }
            }
        } else {
            std::vector<uint64_t> keys(N);
            for (int i = 0; i < N; ++i) keys[i] = bkey(cell(i, D, 0), cell(i, D, 1), cell(i, D, 2));
            std::sort(keys.begin(), keys.end());
            occupied = size_t(std::unique(keys.begin(), keys.end()) - keys.begin());
        }
        if (double(N) / double(occupied) <= leaf_target) break;
    }
    pl->D = D;
    {
        static std::atomic<unsigned long long> next_uid{1};
        pl->uid = next_uid++;
    }
    pl->edge = edge;
    FmmPlanStage& st = *pl->stage;
    for (int k = 0; k < 3; ++k) st.x0[k] = x0[k];
    st.cx.resize(N);
    st.cy.resize(N);
    st.cz.resize(N);
    for (int i = 0; i < N; ++i) {
        st.cx[i] = cell(i, D, 0);
        st.cy[i] = cell(i, D, 1);
        st.cz[i] = cell(i, D, 2);
    }
    // the packed leaf cell of every patch: the leaf-pass neighbourhood test
    // of the near coefficients, which can start before the plan is finished
    if (D <= FMM_CELL_BITS) {
        st.ck.resize(N);
        for (int i = 0; i < N; ++i)
            st.ck[i] = st.cx[i] | (st.cy[i] << FMM_CELL_BITS) | (st.cz[i] << (2 * FMM_CELL_BITS));
        upload(arena, "cellkey", pl->cellkey, st.ck, s);
    }
    return pl;
}

void fmm_plan_finish(FmmPlan* pl, hipStream_t s) {
    if (!pl->stage) return;
    DevArena& arena = *pl->arena;
    const int N = pl->N, D = pl->D;
    const int NP = pl->P * pl->P * pl->P;
    const int order = pl->P;
    const double edge = pl->edge;
    const double* x0 = pl->stage->x0;
    const std::vector<uint32_t>& cx = pl->stage->cx;
    const std::vector<uint32_t>& cy = pl->stage->cy;
    const std::vector<uint32_t>& cz = pl->stage->cz;
    // sort patches by leaf key (Morton-interleaved for locality)
    auto morton = [&](uint32_t x, uint32_t y, uint32_t z) {
        uint64_t m = 0;
        for (int b = 0; b < 21; ++b)
            m |= (uint64_t((x >> b) & 1) << (3 * b + 2)) | (uint64_t((y >> b) & 1) << (3 * b + 1)) |
                 (uint64_t((z >> b) & 1) << (3 * b));
        return m;
    };
    std::vector<uint64_t> code(N);
    for (int i = 0; i < N; ++i) code[i] = morton(cx[i], cy[i], cz[i]);
    std::vector<int> perm(N);
    for (int i = 0; i < N; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return code[a] < code[b]; });
    // boxes per level; leaf level first
    struct Box { uint32_t x, y, z; };
    std::vector<std::vector<Box>> lev(D + 1);
    std::vector<BoxIndex> idx;
    for (int l = 0; l <= D; ++l) idx.emplace_back(l);
    std::vector<int> leaf_src;
    for (int k = 0; k < N; ++k) {
        const int i = perm[k];
        if (idx[D].insert(cx[i], cy[i], cz[i], int(lev[D].size()))) {
            lev[D].push_back({cx[i], cy[i], cz[i]});
            leaf_src.push_back(k);
        }
    }
    leaf_src.push_back(N);
    for (int l = D - 1; l >= 0; --l)
        for (const Box& b : lev[l + 1]) {
            const Box pb{b.x >> 1, b.y >> 1, b.z >> 1};
            if (idx[l].insert(pb.x, pb.y, pb.z, int(lev[l].size()))) lev[l].push_back(pb);
        }
    // global box numbering: level 0 first
    pl->lvl_first.assign(D + 2, 0);
    for (int l = 0; l <= D; ++l) pl->lvl_first[l + 1] = pl->lvl_first[l] + int(lev[l].size());
    const int nbox = pl->lvl_first[D + 1];
    std::vector<double4> box_c(nbox);
    std::vector<double> box_h(D + 1);
    std::vector<int> box_parent(nbox, -1), box_oct(nbox, 0);
    std::vector<std::vector<std::pair<int, int>>> kids(nbox);
    for (int l = 0; l <= D; ++l) {
        const double h = edge / double(1L << l);
        box_h[l] = h;
        for (size_t q = 0; q < lev[l].size(); ++q) {
            const Box& b = lev[l][q];
            const int g = pl->lvl_first[l] + int(q);
            box_c[g] = make_double4(x0[0] + (b.x + 0.5) * h, x0[1] + (b.y + 0.5) * h,
                                    x0[2] + (b.z + 0.5) * h, h);
            if (l > 0) {
                const int par = pl->lvl_first[l - 1] + idx[l - 1].find(b.x >> 1, b.y >> 1, b.z >> 1);
                const int oct = int(b.x & 1) | (int(b.y & 1) << 1) | (int(b.z & 1) << 2);
                box_parent[g] = par;
                box_oct[g] = oct;
                kids[par].push_back({g, oct});
            }
        }
    }
    std::vector<int> child_ptr(nbox + 1, 0), child_box, child_oct;
    for (int g = 0; g < nbox; ++g) {
        for (auto& kc : kids[g]) { child_box.push_back(kc.first); child_oct.push_back(kc.second); }
        child_ptr[g + 1] = int(child_box.size());
    }
    // leaf neighbour lists (27 boxes incl. self)
    const int nleaf = int(lev[D].size());
    std::vector<int> nb_ptr(nleaf + 1, 0), nb_leaf;
    for (int q = 0; q < nleaf; ++q) {
        const Box& b = lev[D][q];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const long x = long(b.x) + dx, y = long(b.y) + dy, z = long(b.z) + dz;
                    const long cells = 1L << D;
                    if (x < 0 || y < 0 || z < 0 || x >= cells || y >= cells || z >= cells) continue;
                    const int q2 = idx[D].find(uint32_t(x), uint32_t(y), uint32_t(z));
                    if (q2 >= 0) nb_leaf.push_back(q2);
                }
        nb_ptr[q + 1] = int(nb_leaf.size());
    }
    for (int q = 0; q < nleaf; ++q) {
        long long srcs = 0;
        for (int k = nb_ptr[q]; k < nb_ptr[q + 1]; ++k)
            srcs += leaf_src[nb_leaf[k] + 1] - leaf_src[nb_leaf[k]];
        pl->n_p2p += srcs * (leaf_src[q + 1] - leaf_src[q]);
    }
    // interaction lists per level (l >= 2), grouped by offset into chunks
    constexpr int G = FmmPlan::CHUNK;
    static_assert(FmmPlan::CHUNK == M2L_G, "plan chunks match the M2L block");
    std::vector<int> pair_src, pair_sym, pair_seg, chunk_seg, seg_tgt;
    std::vector<int4> chunks;
    int n_segs = 0;
    pl->lvl_chunk.assign(D + 2, 0);
    for (int l = 0; l <= D; ++l) {
        pl->lvl_chunk[l] = int(chunks.size());
        if (l < 2) continue;
        std::vector<std::vector<std::array<int, 3>>> byoff(FmmPlan::NCANON);
        for (size_t q = 0; q < lev[l].size(); ++q) {
            const Box& b = lev[l][q];
            const long px = b.x >> 1, py = b.y >> 1, pz = b.z >> 1;
            const long pcells = 1L << (l - 1);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const long X = px + dx, Y = py + dy, Z = pz + dz;
                        if (X < 0 || Y < 0 || Z < 0 || X >= pcells || Y >= pcells || Z >= pcells) continue;
                        const int pq = idx[l - 1].find(uint32_t(X), uint32_t(Y), uint32_t(Z));
                        if (pq < 0) continue;
                        const int pg = pl->lvl_first[l - 1] + pq;
                        for (auto& kc : kids[pg]) {
                            const int sg = kc.first;
                            const Box& sb = lev[l][sg - pl->lvl_first[l]];
                            const int tx = int(long(sb.x) - long(b.x)), ty = int(long(sb.y) - long(b.y)),
                                      tz = int(long(sb.z) - long(b.z));
                            if (std::max(std::abs(tx), std::max(std::abs(ty), std::abs(tz))) <= 1) continue;
                            const OffsetClass& oc =
                                offset_classes()[size_t(((tx + 3) * 7 + (ty + 3)) * 7 + (tz + 3))];
                            byoff[oc.canon].push_back({pl->lvl_first[l] + int(q), sg, oc.sym});
                        }
                    }
        }
        for (int off = 0; off < FmmPlan::NCANON; ++off) {
            auto& v = byoff[off];
            for (size_t k = 0; k < v.size(); k += G) {
                const int b0 = int(pair_src.size());
                chunk_seg.push_back(n_segs);
                for (size_t j = k; j < std::min(v.size(), k + G); ++j) {
                    // a segment: a run of one target's pairs inside the chunk
                    // (pairs of a class are listed target by target)
                    if (j == k || v[j][0] != v[j - 1][0]) {
                        seg_tgt.push_back(v[j][0]);
                        ++n_segs;
                    }
                    pair_src.push_back(v[j][1]);
                    pair_sym.push_back(v[j][2]);
                    pair_seg.push_back(n_segs - 1);
                }
                chunks.push_back(make_int4(off, b0, int(pair_src.size()), l));
            }
        }
    }
    pl->lvl_chunk[D + 1] = int(chunks.size());
    chunk_seg.push_back(n_segs);
    if (D >= 2 && pl->lvl_chunk[2] != 0) throw std::runtime_error("M2L chunks must start at level 2");
    pl->n_pairs = int(pair_src.size());
    pl->n_segs = n_segs;
    // per target box its segments, in increasing order (the order k_m2l_sum
    // adds them in)
    std::vector<int> tgt_ptr(nbox + 1, 0), tgt_seg(static_cast<size_t>(n_segs));
    for (int tg : seg_tgt) tgt_ptr[tg + 1]++;
    for (int g = 0; g < nbox; ++g) tgt_ptr[g + 1] += tgt_ptr[g];
    {
        std::vector<int> fill(tgt_ptr.begin(), tgt_ptr.end() - 1);
        for (int k = 0; k < n_segs; ++k) tgt_seg[fill[seg_tgt[k]]++] = k;
    }
    if (pl->engine == 1) {
        // children of every box (by octant), the X2X lists of every box at
        // levels >= 2 (per direction, sources in the order met) and the L2L
        // rows of every level grouped by child octant
        std::vector<int> child8(size_t(nbox) * 8, -1);
        for (int g = 0; g < nbox; ++g)
            for (auto& kc : kids[g]) child8[size_t(g) * 8 + kc.second] = kc.first;
        const int base = D >= 2 ? pl->lvl_first[2] : nbox;
        std::vector<int> xl_ptr(size_t(nbox - base) * 6 + 1, 0), xl_src, xl_code;
        double rot[6][3][3];
        pw_direction_frames(rot);
        for (int l = 2; l <= D; ++l) {
            const long pcells = 1L << (l - 1);
            for (size_t q = 0; q < lev[l].size(); ++q) {
                const Box& b = lev[l][q];
                std::vector<std::pair<int, int>> byd[6];
                const long px = b.x >> 1, py = b.y >> 1, pz = b.z >> 1;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const long X = px + dx, Y = py + dy, Z = pz + dz;
                            if (X < 0 || Y < 0 || Z < 0 || X >= pcells || Y >= pcells || Z >= pcells) continue;
                            const int pq = idx[l - 1].find(uint32_t(X), uint32_t(Y), uint32_t(Z));
                            if (pq < 0) continue;
                            for (auto& kc : kids[pl->lvl_first[l - 1] + pq]) {
                                const Box& sb = lev[l][kc.first - pl->lvl_first[l]];
                                // target centre - source centre, box units
                                const int tx = int(long(b.x) - long(sb.x)), ty = int(long(b.y) - long(sb.y)),
                                          tz = int(long(b.z) - long(sb.z));
                                const int dir = pw_direction(tx, ty, tz);
                                if (dir < 0) continue;
                                int tl[3];
                                for (int a = 0; a < 3; ++a)
                                    tl[a] = int(std::lround(rot[dir][0][a] * tx + rot[dir][1][a] * ty +
                                                            rot[dir][2][a] * tz));
                                if (tl[2] < 2 || tl[2] > 3 || std::abs(tl[0]) > 3 || std::abs(tl[1]) > 3)
                                    throw std::runtime_error("plane-wave list: offset out of range");
                                byd[dir].push_back({kc.first, (tl[2] - 2) * 49 + (tl[1] + 3) * 7 + (tl[0] + 3)});
                            }
                        }
                const size_t row = size_t(pl->lvl_first[l] + int(q) - base) * 6;
                for (int dir = 0; dir < 6; ++dir) {
                    for (auto& e : byd[dir]) {
                        xl_src.push_back(e.first);
                        xl_code.push_back(e.second);
                    }
                    xl_ptr[row + dir + 1] = int(xl_src.size());
                }
            }
        }
        pl->pw_n_x2x = (long long)xl_src.size();
        std::vector<int> l2l_rows, l2l_gptr(size_t(D + 1) * 9, 0);
        for (int l = 0; l <= D; ++l) {
            for (int o = 0; o < 8; ++o) {
                l2l_gptr[size_t(l) * 9 + o] = int(l2l_rows.size());
                if (l >= 3)
                    for (int g = pl->lvl_first[l]; g < pl->lvl_first[l + 1]; ++g)
                        if (box_oct[g] == o) l2l_rows.push_back(g);
            }
            l2l_gptr[size_t(l) * 9 + 8] = int(l2l_rows.size());
        }
        if (l2l_rows.empty()) l2l_rows.push_back(0);
        if (xl_src.empty()) {
            xl_src.push_back(0);
            xl_code.push_back(0);
        }
        upload(arena, "pw_child8", pl->pw_child8, child8, s);
        upload(arena, "pw_xl_ptr", pl->pw_xl_ptr, xl_ptr, s);
        upload(arena, "pw_xl_src", pl->pw_xl_src, xl_src, s);
        upload(arena, "pw_xl_code", pl->pw_xl_code, xl_code, s);
        upload(arena, "pw_l2l_rows", pl->pw_l2l_rows, l2l_rows, s);
        upload(arena, "pw_l2l_gptr", pl->pw_l2l_gptr, l2l_gptr, s);
        pl->pw_l2l_gptr_h = l2l_gptr;
        std::vector<int> box_level(nbox, 0);
        for (int l = 0; l <= D; ++l)
            for (int g = pl->lvl_first[l]; g < pl->lvl_first[l + 1]; ++g) box_level[g] = l;
        upload(arena, "pw_box_level", pl->pw_box_level, box_level, s);
        // the host vectors must live until the uploads complete
        hip_check(hipStreamSynchronize(s), "sync plane-wave lists");
    }
    // 1-D Chebyshev tables
    const int P = order;
    std::vector<double> xi(P), tn(P * P), t_oct(2 * P * P);
    for (int k = 0; k < P; ++k) xi[k] = std::cos((2.0 * k + 1.0) * PI / (2.0 * P));
    for (int k = 0; k < P; ++k)
        for (int j = 0; j < P; ++j) tn[k * P + j] = std::cos(j * (2.0 * k + 1.0) * PI / (2.0 * P));
    auto S1 = [&](int k, double u) {
        double a = 0, tm2 = 1, tm1 = u;
        for (int j = 1; j < P; ++j) {
            const double tj = (j == 1) ? u : 2 * u * tm1 - tm2;
            if (j >= 2) { tm2 = tm1; tm1 = tj; }
            a += tn[k * P + j] * tj;
        }
        return 1.0 / P + 2.0 / P * a;
    };
    // t_oct[o][a][a2] = S_P(xi_a, (xi_a2 + (o ? 1 : -1)) / 2)
    for (int o = 0; o < 2; ++o)
        for (int a = 0; a < P; ++a)
            for (int a2 = 0; a2 < P; ++a2)
                t_oct[(o * P + a) * P + a2] = S1(a, (xi[a2] + (o ? 1.0 : -1.0)) / 2.0);
    pl->box_h = box_h;
    pl->edge = edge;
    pl->nbox = nbox;
    pl->nleaf = nleaf;
    pl->dv.D = D;
    pl->dv.leaf_first = pl->lvl_first[D];
    upload(arena, "box_c", pl->dv.box_c, box_c, s);
    upload(arena, "box_h", pl->dv.box_h, box_h, s);
    upload(arena, "box_parent", pl->dv.box_parent, box_parent, s);
    upload(arena, "box_oct", pl->dv.box_oct, box_oct, s);
    upload(arena, "child_ptr", pl->dv.child_ptr, child_ptr, s);
    upload(arena, "child_box", pl->dv.child_box, child_box, s);
    upload(arena, "child_oct", pl->dv.child_oct, child_oct, s);
    upload(arena, "leaf_src", pl->dv.leaf_src, leaf_src, s);
    upload(arena, "perm", pl->dv.perm, perm, s);
    // patch -> FMM order and FMM order -> leaf box; the host copies live
    // until the stream is synchronised below
    std::vector<int> iperm(N), tleaf(N);
    for (int k = 0; k < N; ++k) iperm[perm[k]] = k;
    for (int q = 0; q < nleaf; ++q)
        for (int k = leaf_src[q]; k < leaf_src[q + 1]; ++k) tleaf[k] = q;
    upload(arena, "iperm", pl->d_iperm, iperm, s);
    upload(arena, "tleaf", pl->d_tleaf, tleaf, s);
    upload(arena, "nb_ptr", pl->dv.nb_ptr, nb_ptr, s);
    upload(arena, "nb_leaf", pl->dv.nb_leaf, nb_leaf, s);
    upload(arena, "pair_src", pl->dv.pair_src, pair_src, s);
    upload(arena, "pair_sym", pl->dv.pair_sym, pair_sym, s);
    upload(arena, "pair_seg", pl->dv.pair_seg, pair_seg, s);
    upload(arena, "chunk_seg", pl->dv.chunk_seg, chunk_seg, s);
    // node maps of the 48 symmetries: psi (canonical -> actual) and phi
    // (actual -> canonical), order^3 entries each (host copies live until the
    // stream is synchronised below)
    const int NPn = order * order * order;
    std::vector<unsigned short> psi(size_t(48) * NPn), phi(size_t(48) * NPn);
    if (pl->engine == 0) {
        for (int k = 0; k < 48; ++k)
            for (int m = 0; m < NPn; ++m) {
                int a = 0;
                switch (order) {
                    case 3: a = node_of<3>(m, symmetry_codes()[k]); break;
                    case 4: a = node_of<4>(m, symmetry_codes()[k]); break;
                    case 5: a = node_of<5>(m, symmetry_codes()[k]); break;
                    default: a = node_of<6>(m, symmetry_codes()[k]); break;
                }
                psi[size_t(k) * NPn + m] = (unsigned short)a;
                phi[size_t(k) * NPn + a] = (unsigned short)m;
            }
        upload(arena, "psi", pl->dv.psi, psi, s);
        upload(arena, "phi", pl->dv.phi, phi, s);
    }
    std::vector<int4> canon(FmmPlan::NCANON);
    {
        for (int c = 0; c < FmmPlan::NCANON; ++c) {
            const auto t = canonical_list()[c];
            canon[c] = make_int4(t[0], t[1], t[2], 0);
        }
        upload(arena, "canon", pl->d_canon, canon, s);
    }
    upload(arena, "tgt_ptr", pl->dv.tgt_ptr, tgt_ptr, s);
    upload(arena, "tgt_seg", pl->dv.tgt_seg, tgt_seg, s);
    upload(arena, "tn", pl->dv.tn, tn, s);
    upload(arena, "t_oct", pl->dv.t_oct, t_oct, s);
    upload(arena, "chunks", pl->d_chunks, chunks, s);
    upload(arena, "xi", pl->d_xi, xi, s);
    pl->ktab64.assign(D + 2, nullptr);
    pl->ktab32.assign(D + 2, nullptr);
    pl->src = arena.get<Src64>("src", size_t(std::max(N, 1)));
    const size_t ne = size_t(nbox) * 3 * NP;
    pl->n_exp = ne;
    hip_check(hipStreamSynchronize(s), "sync fmm plan");
    delete pl->stage;  // the cell keys are uploaded too (same stream)
    pl->stage = nullptr;
}

double fmm_algorithmic_bytes(const FmmPlan* pl, int fp32) {
    // every far-field array read once and written once per mat-vec
    const double T = fp32 ? 4.0 : 8.0;
    const double N = pl->N;
    const double ncoef = pl->engine == 1 ? double(pl->pw_jp) : double(pl->P) * pl->P * pl->P;
    const double E = 3.0 * ncoef * T;  // one box's three expansions
    const double nbox = pl->nbox, nleaf = pl->nleaf;
    double b = N * (16.0 + 64.0 + 64.0);  // pack: x and patches in, sources out
    b += N * 64.0 + nleaf * E;            // P2M
    b += 2.0 * (nbox - 1.0) * E;          // M2M: children read, parents written
    b += 2.0 * (nbox - 1.0) * E;          // translations: source moments read, locals written
    b += 3.0 * (nbox - nleaf) * E;        // L2L: parents read, children read and written
    b += nleaf * E + N * 64.0 + 4.0 * N * 8.0;  // L2P: locals, patches, far vector in and out
    if (pl->engine == 1 && pl->D >= 2) {
        // plane waves: M2X writes W, X2X reads it and writes Wacc, X2L reads it
        const double Wb = 3.0 * pl->pw_nwp * T;
        b += 4.0 * (nbox - pl->lvl_first[2]) * Wb;
    }
    // leaf pass: its pair list (one index per pair), sources, y in and out
    const double pairs = pl->n_res >= 0 ? double(pl->n_res) : double(pl->n_p2p);
    b += pairs * 4.0 + N * 64.0 + 4.0 * N * 8.0;
    return b;
}

FmmPlan* fmm_plan_create(int N, const double* centroid, int order, int leaf_target,
                         DevArena& arena, hipStream_t s, int engine) {
    FmmPlan* pl = fmm_plan_begin(N, centroid, order, leaf_target, arena, s, engine);
    fmm_plan_finish(pl, s);
    return pl;
}

void fmm_plan_destroy(FmmPlan* p) {
    if (!p) return;
    delete p->stage;
    delete p;
}

template <typename T, int P>
static void make_tables(FmmPlan* pl, double kappa, std::vector<T*>& tabs, T**& ytab,
                        hipStream_t s) {
    constexpr int NP = P * P * P;
    const size_t n = size_t(FmmPlan::NCANON) * NP * NP;
    const int blocks = int((n + 255) / 256);
    const std::string pre = sizeof(T) == 8 ? "ktab64_" : "ktab32_";
    tabs.assign(pl->D + 2, nullptr);
    // slot D+1: Laplace (level independent, h = 1); slot l: Yukawa at level l
    tabs[pl->D + 1] = pl->arena->get<T>(pre + "laplace", n);
    hipLaunchKernelGGL((k_ktab<T, P>), dim3(blocks), dim3(256), 0, s, pl->d_xi, pl->d_canon,
                       FmmPlan::NCANON, 1.0, 0.0, 1, tabs[pl->D + 1]);
    for (int l = 2; l <= pl->D; ++l) {
        tabs[l] = pl->arena->get<T>(pre + std::to_string(l), n);
        hipLaunchKernelGGL((k_ktab<T, P>), dim3(blocks), dim3(256), 0, s, pl->d_xi, pl->d_canon,
                           FmmPlan::NCANON, pl->box_h[l], kappa, 0, tabs[l]);
    }
    hip_check(hipGetLastError(), "k_ktab launch");
    ytab = pl->arena->get<T*>(pre + "ptrs", tabs.size());
    hip_check(hipMemcpyAsync(ytab, tabs.data(), tabs.size() * sizeof(T*), hipMemcpyHostToDevice, s),
              "H2D ytab");
}

// M2L result slots: 3 x P^3 values per target segment
static size_t slot_count(const FmmPlan* pl) {
    const int NP = pl->P * pl->P * pl->P;
    return std::max<size_t>(1, size_t(pl->n_segs) * 3 * NP);
}

void fmm_prepare(FmmPlan* pl, double kappa, bool fp64, bool fp32, hipStream_t s) {
    if (pl->tables_kappa == kappa && (!fp64 || pl->have64) && (!fp32 || pl->have32)) return;
    if (pl->tables_kappa != kappa) {
        pl->have64 = pl->have32 = false;
        pl->tables_kappa = kappa;
    }
    if (pl->engine == 1) {
        pw_prepare(pl, kappa, fp64, fp32, s);
        hip_check(hipStreamSynchronize(s), "sync plane-wave tables");
        return;
    }
    if (fp64 && !pl->have64) {
        switch (pl->P) {
            case 3: make_tables<double, 3>(pl, kappa, pl->ktab64, pl->ytab64, s); break;
            case 4: make_tables<double, 4>(pl, kappa, pl->ktab64, pl->ytab64, s); break;
            case 5: make_tables<double, 5>(pl, kappa, pl->ktab64, pl->ytab64, s); break;
            case 6: make_tables<double, 6>(pl, kappa, pl->ktab64, pl->ytab64, s); break;
        }
        const size_t ns = slot_count(pl);
        pl->exp64 = pl->arena->get<double>("exp64", 2 * pl->n_exp);
        pl->slots64 = pl->arena->get<double>("slots64", ns);
        pl->have64 = true;
    }
    if (fp32 && !pl->have32) {
        switch (pl->P) {
            case 3: make_tables<float, 3>(pl, kappa, pl->ktab32, pl->ytab32, s); break;
            case 4: make_tables<float, 4>(pl, kappa, pl->ktab32, pl->ytab32, s); break;
            case 5: make_tables<float, 5>(pl, kappa, pl->ktab32, pl->ytab32, s); break;
            case 6: make_tables<float, 6>(pl, kappa, pl->ktab32, pl->ytab32, s); break;
        }
        const size_t ns = slot_count(pl);
        pl->exp32 = pl->arena->get<float>("exp32", 2 * pl->n_exp);
        pl->slots32 = pl->arena->get<float>("slots32", ns);
        pl->have32 = true;
    }
    hip_check(hipStreamSynchronize(s), "sync fmm tables");
}

template <typename T, int P>
static void evaluate(FmmPlan* pl, hipStream_t s, const Patch64* patches, const double* x,
                     double kappa, T* M, T* L, const std::vector<T*>& tabs, T* const* ytab,
                     T* slots, double* y, hipEvent_t* m2l_ev, hipEvent_t pack_ev) {
    constexpr int NP = P * P * P;
    constexpr int NT = ((NP + 63) / 64) * 64;
    const FmmDev& d = pl->dv;
    launch_far_pack(s, pl->N, patches, d.perm, x, pl->src);
    if (pack_ev) hip_check(hipEventRecord(pack_ev, s), "hipEventRecord");
    {
        const size_t n = pl->n_exp;
        hipLaunchKernelGGL((k_zero<T>), dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, L);
    }
    const int D = pl->D;
    hipLaunchKernelGGL((k_p2m<T, P>), dim3(pl->nleaf), dim3(NT), 0, s, d, pl->src, M);
    // parents at levels D-1 .. 2 (level-1 multipoles are never translated)
    for (int l = D; l >= 3; --l) {
        const int first = pl->lvl_first[l - 1], cnt = pl->lvl_first[l] - first;
        if (cnt > 0) hipLaunchKernelGGL((k_m2m<T, P>), dim3(cnt), dim3(256), 0, s, d, first, M);
    }
    {
        // chunks exist from level 2 on and are numbered from 0; their
        // segments (slots) are numbered in chunk order
        const int nch = D >= 2 ? pl->lvl_chunk[D + 1] : 0;
        if (m2l_ev) hip_check(hipEventRecord(m2l_ev[0], s), "hipEventRecord");
        if (nch > 0) {
            hipLaunchKernelGGL((k_m2l<T, P>), dim3(3 * nch), dim3(512), 0, s, d, pl->d_chunks,
                               tabs[D + 1], ytab, M, slots);
            const int b0 = pl->lvl_first[2], nb = pl->lvl_first[D + 1] - b0;
            hipLaunchKernelGGL((k_m2l_sum<T, P>), dim3(nb), dim3(256), 0, s, d, b0, slots, L);
        }
        if (m2l_ev) hip_check(hipEventRecord(m2l_ev[1], s), "hipEventRecord");
    }
    for (int l = 2; l < D; ++l) {
        const int first = pl->lvl_first[l + 1], cnt = pl->lvl_first[l + 2] - first;
        if (cnt > 0) hipLaunchKernelGGL((k_l2l<T, P>), dim3(cnt), dim3(256), 0, s, d, first, L);
    }
    if (D >= 2)
        hipLaunchKernelGGL((k_l2p<T, P>), dim3(pl->nleaf), dim3(64), 0, s, d, patches, L, y,
                           pl->N);
    hip_check(hipGetLastError(), "fmm launch");
}

int fmm_canonical(int tx, int ty, int tz, int order, int* canon_xyz, int* node_map) {
    int sym = 0;
    const int c = canonical_offset(tx, ty, tz, &sym);
    const auto t = canonical_list()[c];
    for (int k = 0; k < 3; ++k) canon_xyz[k] = t[k];
    const int NP = order * order * order;
    for (int m = 0; m < NP; ++m) {
        switch (order) {
            case 3: node_map[m] = node_of<3>(m, sym); break;
            case 4: node_map[m] = node_of<4>(m, sym); break;
            case 5: node_map[m] = node_of<5>(m, sym); break;
            case 6: node_map[m] = node_of<6>(m, sym); break;
            default: throw std::runtime_error("FMM order must be 3..6");
        }
    }
    return c;
}

void fmm_p2p(FmmPlan* pl, hipStream_t s, int fp32, const Patch64* patches, double kappa,
             double* y) {
    if (pl->n_res >= 0) {
        // the near coefficients hold the explicit pairs of the leaf pass in full
        if (pl->n_res > 0) launch_p2p_residual(pl, s, fp32, patches, kappa, y);
        return;
    }
    if (fp32)
        hipLaunchKernelGGL((k_p2p<float>), dim3(pl->nleaf), dim3(256), 0, s, pl->dv, pl->src,
                           patches, float(kappa), y, pl->N);
    else
        hipLaunchKernelGGL((k_p2p<double>), dim3(pl->nleaf), dim3(256), 0, s, pl->dv, pl->src,
                           patches, kappa, y, pl->N);
    hip_check(hipGetLastError(), "k_p2p launch");
}

void fmm_evaluate(FmmPlan* pl, hipStream_t s, int fp32, const Patch64* patches, const double* x,
                  double kappa, double* y, hipEvent_t* m2l_ev, hipEvent_t pack_ev) {
    if (pl->engine == 1) {
        launch_far_pack(s, pl->N, patches, pl->dv.perm, x, pl->src);
        if (pack_ev) hip_check(hipEventRecord(pack_ev, s), "hipEventRecord");
        pw_evaluate(pl, s, fp32, patches, kappa, y, m2l_ev);
        return;
    }
    if (fp32) {
        if (!pl->have32) throw std::runtime_error("FMM FP32 tables not prepared");
        float* M = pl->exp32;
        float* L = pl->exp32 + pl->n_exp;
        switch (pl->P) {
            case 3: evaluate<float, 3>(pl, s, patches, x, kappa, M, L, pl->ktab32, pl->ytab32, pl->slots32, y, m2l_ev, pack_ev); break;
            case 4: evaluate<float, 4>(pl, s, patches, x, kappa, M, L, pl->ktab32, pl->ytab32, pl->slots32, y, m2l_ev, pack_ev); break;
            case 5: evaluate<float, 5>(pl, s, patches, x, kappa, M, L, pl->ktab32, pl->ytab32, pl->slots32, y, m2l_ev, pack_ev); break;
            case 6: evaluate<float, 6>(pl, s, patches, x, kappa, M, L, pl->ktab32, pl->ytab32, pl->slots32, y, m2l_ev, pack_ev); break;
        }
    } else {
        if (!pl->have64) throw std::runtime_error("FMM FP64 tables not prepared");
        double* M = pl->exp64;
        double* L = pl->exp64 + pl->n_exp;
        switch (pl->P) {
            case 3: evaluate<double, 3>(pl, s, patches, x, kappa, M, L, pl->ktab64, pl->ytab64, pl->slots64, y, m2l_ev, pack_ev); break;
            case 4: evaluate<double, 4>(pl, s, patches, x, kappa, M, L, pl->ktab64, pl->ytab64, pl->slots64, y, m2l_ev, pack_ev); break;
            case 5: evaluate<double, 5>(pl, s, patches, x, kappa, M, L, pl->ktab64, pl->ytab64, pl->slots64, y, m2l_ev, pack_ev); break;
            case 6: evaluate<double, 6>(pl, s, patches, x, kappa, M, L, pl->ktab64, pl->ytab64, pl->slots64, y, m2l_ev, pack_ev); break;
        }
    }
}

}  // namespace beepdev
