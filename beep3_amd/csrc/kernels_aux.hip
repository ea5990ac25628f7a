// Right-hand side, energy coefficients and the FP64 vector kernels of the
// device GMRES driver.
#include "device.hpp"
#include "curved.hpp"

namespace beepdev {

// Coulomb potential and gradient (field = +grad phi, src/fmm/eval_pt.h:175-197)
// of all charges at a point, FP64.  Charges are staged through LDS in tiles.
constexpr int CH_TILE = 256;

// f_rhs[i] = scale * sum_q w_q phi(x_q);  h_rhs[i] = scale * sum_q w_q grad phi . n_q
// (BEEP::set_rhs, beep.cpp:1124-1260).  One thread per patch.
__global__ __launch_bounds__(256) void k_rhs(int n, const int* __restrict__ qual_ptr,
                                             const Pt64* __restrict__ qual,
                                             const Patch64* __restrict__ patches, int nch,
                                             const double* __restrict__ xyzq, double scale,
                                             double* __restrict__ b) {
    constexpr int QB = 8;  // qualocation points in flight per thread
    __shared__ double4 ch[CH_TILE];
    __shared__ int max_chunks;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    const int q0 = valid ? qual_ptr[i] : 0, q1 = valid ? qual_ptr[i + 1] : 0;
    if (threadIdx.x == 0) max_chunks = 0;
    __syncthreads();
    atomicMax(&max_chunks, (q1 - q0 + QB - 1) / QB);
    __syncthreads();
    const int nchunks = max_chunks;  // same trip count for every thread of the block
    const Patch64 p = valid ? patches[i] : Patch64{0, 0, 0, 0, 0, 0, 0, 0};
    double fr = 0, hr = 0;
    for (int chk = 0; chk < nchunks; ++chk) {
        const int qb = q0 + chk * QB;
        const int nq = max(0, min(QB, q1 - qb));
        double px[QB], py[QB], pz[QB], pot[QB], gx[QB], gy[QB], gz[QB];
        for (int a = 0; a < QB; ++a) {
            pot[a] = gx[a] = gy[a] = gz[a] = 0;
            px[a] = py[a] = pz[a] = 0;
            if (a < nq) {
                const Pt64 q = qual[qb + a];
                px[a] = p.cx + q.x; py[a] = p.cy + q.y; pz[a] = p.cz + q.z;
            }
        }
        for (int c0 = 0; c0 < nch; c0 += CH_TILE) {
            __syncthreads();
            if (c0 + threadIdx.x < nch) {
                const double* c = xyzq + 4 * (c0 + threadIdx.x);
                ch[threadIdx.x] = make_double4(c[0], c[1], c[2], c[3]);
            }
            __syncthreads();
            const int m = min(CH_TILE, nch - c0);
            for (int a = 0; a < nq; ++a) {
                double pa = 0, xa = 0, ya = 0, za = 0;
                for (int k = 0; k < m; ++k) {
                    const double4 c = ch[k];
                    const double rx = px[a] - c.x, ry = py[a] - c.y, rz = pz[a] - c.z;
                    const double d2 = rx * rx + ry * ry + rz * rz;
                    // a charge on the point contributes nothing (id = 0)
                    const double id = d2 <= 2.220446049250313e-16 ? 0.0 : rsqrt_nr(d2);
                    const double pt = c.w * id;
                    pa += pt;
                    const double ft = -pt * id * id;
                    xa += ft * rx; ya += ft * ry; za += ft * rz;
                }
                pot[a] += pa; gx[a] += xa; gy[a] += ya; gz[a] += za;
            }
        }
        for (int a = 0; a < nq; ++a) {
            const Pt64 q = qual[qb + a];
            fr += q.w * pot[a];
            hr += q.w * (gx[a] * q.nx + gy[a] * q.ny + gz[a] * q.nz);
        }
    }
    if (valid) {
        b[i] = fr * scale;
        b[n + i] = hr * scale;
    }
}

void launch_rhs(hipStream_t s, int n, const int* qual_ptr, const Pt64* qual,
                const Patch64* patches, int n_charges, const double* xyzq, double scale,
                double* b) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_rhs, dim3((n + 255) / 256), dim3(256), 0, s, n, qual_ptr, qual, patches,
                       n_charges, xyzq, scale, b);
    hip_check(hipGetLastError(), "k_rhs launch");
}

// Energy coefficients (Mesh::init_energy_precalcs, mesh.cpp:730-887) before
// the Gauss-law renormalisation: over the galerkin points of each patch,
//   ec_f = sum w (grad phi . n) / 4pi,  ec_h = sum w phi / 4pi,
// with phi from the charges of the patch's own instance.  One wave per patch,
// lane = charge (a patch has a few dozen galerkin points and an instance
// thousands of charges).  The points, as position and w n, are staged in
// LDS (read by every lane alike); each lane takes two charges per pass so a
// point read serves two pairs.
constexpr int EC_PTS = 64;  // galerkin points staged per pass

__global__ __launch_bounds__(64) void k_energy_coeffs(int n, const int* __restrict__ gal_ptr,
                                                      const Pt64* __restrict__ gal,
                                                      const Patch64* __restrict__ patches,
                                                      const int* __restrict__ patch_inst,
                                                      const int* __restrict__ inst_charge_ptr,
                                                      const double* __restrict__ xyzq,
                                                      double* __restrict__ ec_f,
                                                      double* __restrict__ ec_h) {
    __shared__ double4 pxw[EC_PTS];  // point position, w
    __shared__ double4 pwn[EC_PTS];  // w n
    const int i = blockIdx.x;
    if (i >= n) return;
    const int t = threadIdx.x;
    const int g0 = gal_ptr[i], ng = gal_ptr[i + 1] - g0;
    const int inst = patch_inst[i];
    const int c0 = inst_charge_ptr[inst], c1 = inst_charge_ptr[inst + 1];
    const Patch64 p = patches[i];
    double ef = 0, eh = 0;
    for (int a0 = 0; a0 < ng; a0 += EC_PTS) {
        const int na = min(EC_PTS, ng - a0);
        __syncthreads();  // previous points consumed
        if (t < na) {
            const Pt64 q = gal[g0 + a0 + t];
            pxw[t] = make_double4(p.cx + q.x, p.cy + q.y, p.cz + q.z, q.w);
            pwn[t] = make_double4(q.w * q.nx, q.w * q.ny, q.w * q.nz, 0.0);
        }
        __syncthreads();
        for (int c = c0 + 2 * t; c < c1; c += 128) {
            // two charges (the second may be past the end: zero charge)
            const bool two = c + 1 < c1;
            const double ax = xyzq[4 * c], ay = xyzq[4 * c + 1], az = xyzq[4 * c + 2];
            const double aq = xyzq[4 * c + 3];
            const double bx = two ? xyzq[4 * c + 4] : ax, by = two ? xyzq[4 * c + 5] : ay;
            const double bz = two ? xyzq[4 * c + 6] : az, bq = two ? xyzq[4 * c + 7] : 0.0;
            for (int k = 0; k < na; ++k) {
                const double4 xw = pxw[k], wn = pwn[k];
                {
                    const double rx = xw.x - ax, ry = xw.y - ay, rz = xw.z - az;
                    const double d2 = rx * rx + ry * ry + rz * rz;
                    // a charge on the point contributes nothing (id = 0)
                    const double id = d2 <= 2.220446049250313e-16 ? 0.0 : rsqrt_nr(d2);
                    const double pt = aq * id;
                    eh += pt * xw.w;
                    ef -= pt * id * id * (rx * wn.x + ry * wn.y + rz * wn.z);
                }
                {
                    const double rx = xw.x - bx, ry = xw.y - by, rz = xw.z - bz;
                    const double d2 = rx * rx + ry * ry + rz * rz;
                    const double id = d2 <= 2.220446049250313e-16 ? 0.0 : rsqrt_nr(d2);
                    const double pt = bq * id;
                    eh += pt * xw.w;
                    ef -= pt * id * id * (rx * wn.x + ry * wn.y + rz * wn.z);
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        ef += __shfl_down(ef, off, 64);
        eh += __shfl_down(eh, off, 64);
    }
    if (t == 0) {
        ec_f[i] = ef * FOUR_PI_INV;
        ec_h[i] = eh * FOUR_PI_INV;
    }
}

void launch_energy_coeffs(hipStream_t s, int n, const int* gal_ptr, const Pt64* gal,
                          const Patch64* patches, const int* patch_inst,
                          const int* inst_charge_ptr, const double* xyzq, double* ec_f,
                          double* ec_h) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_energy_coeffs, dim3(n), dim3(64), 0, s, n, gal_ptr, gal, patches,
                       patch_inst, inst_charge_ptr, xyzq, ec_f, ec_h);
    hip_check(hipGetLastError(), "k_energy_coeffs launch");
}

// ---- FP64 vector kernels for GMRES ----
// partial[q * nb + blk] = sum over the block's rows of V[q*n + k] * w[k]
__global__ __launch_bounds__(256) void k_multi_dot(int n, const double* __restrict__ V, int ncols,
                                                   const double* __restrict__ w,
                                                   double* __restrict__ partial) {
    __shared__ double red[256];
    const int q = blockIdx.y;
    const int nb = gridDim.x;
    double acc = 0;
    const double* v = V + size_t(q) * n;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += nb * 256) acc += v[k] * w[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[size_t(q) * nb + blockIdx.x] = red[0];
}

void launch_multi_dot(hipStream_t s, int n, const double* V, int ncols, const double* w,
                      double* partial, int n_part_blocks) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL(k_multi_dot, dim3(n_part_blocks, ncols), dim3(256), 0, s, n, V, ncols, w,
                       partial);
    hip_check(hipGetLastError(), "k_multi_dot launch");
}

__global__ __launch_bounds__(256) void k_reduce_partials(int ncols, const double* __restrict__ partial,
                                                         int nb, double* __restrict__ out) {
    __shared__ double red[256];
    const int q = blockIdx.x;
    double acc = 0;
    for (int k = threadIdx.x; k < nb; k += 256) acc += partial[size_t(q) * nb + k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

void launch_reduce_partials(hipStream_t s, int ncols, const double* partial, int n_part_blocks,
                            double* out) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL(k_reduce_partials, dim3(ncols), dim3(256), 0, s, ncols, partial,
                       n_part_blocks, out);
    hip_check(hipGetLastError(), "k_reduce_partials launch");
}

// w[k] -= sum_q h[q] V[q*n + k]
__global__ __launch_bounds__(256) void k_gemv_sub(int n, const double* __restrict__ V, int ncols,
                                                  const double* __restrict__ h,
                                                  double* __restrict__ w) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    double acc = 0;
    for (int q = 0; q < ncols; ++q) acc += h[q] * V[size_t(q) * n + k];
    w[k] -= acc;
}

void launch_gemv_sub(hipStream_t s, int n, const double* V, int ncols, const double* h,
                     double* w) {
    if (n <= 0 || ncols <= 0) return;
    hipLaunchKernelGGL(k_gemv_sub, dim3((n + 255) / 256), dim3(256), 0, s, n, V, ncols, h, w);
    hip_check(hipGetLastError(), "k_gemv_sub launch");
}

// y = a x + b y
__global__ void k_axpby(int n, double a, const double* __restrict__ x, double b,
                        double* __restrict__ y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = a * x[k] + b * y[k];
}

void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_axpby, dim3((n + 255) / 256), dim3(256), 0, s, n, a, x, b, y);
    hip_check(hipGetLastError(), "k_axpby launch");
}

// y = x * (scale_ptr ? 1 / *scale_ptr : scale); a zero divisor is replaced by
// the smallest normal double (the NEAR0 guard of beep.cpp:1366)
__global__ void k_scale_copy(int n, const double* __restrict__ x, const double* __restrict__ sp,
                             double scale, double* __restrict__ y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double sc = scale;
    if (sp) {
        const double v = *sp;
        sc = 1.0 / (v == 0 ? 2.2250738585072014e-308 : v);
    }
    if (k < n) y[k] = x[k] * sc;
}

void launch_scale_copy(hipStream_t s, int n, const double* x, const double* scale_ptr,
                       double scale, double* y) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scale_copy, dim3((n + 255) / 256), dim3(256), 0, s, n, x, scale_ptr,
                       scale, y);
    hip_check(hipGetLastError(), "k_scale_copy launch");
}

__global__ void k_fill(int n, double v, double* __restrict__ y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = v;
}

void launch_fill(hipStream_t s, int n, double v, double* y) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, n, v, y);
    hip_check(hipGetLastError(), "k_fill launch");
}

// ---- Arnoldi step with the step index j read from device memory, so that
// one captured HIP graph serves every step of a GMRES cycle ----
__global__ void k_set_int(int* __restrict__ p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_inc_int(int* __restrict__ p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += 1;
}

// out = V[j]
__global__ void k_copy_col(int n, const double* __restrict__ V, const int* __restrict__ jp,
                           double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = V[size_t(*jp) * n + k];
}

// partial[q * nb + blk] = block sum of V[q] . w for q <= j (grid y = m + 1)
__global__ __launch_bounds__(256) void k_multi_dot_j(int n, const double* __restrict__ V,
                                                     const int* __restrict__ jp,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ partial) {
    const int q = blockIdx.y;
    if (q > *jp) return;
    __shared__ double red[256];
    const int nb = gridDim.x;
    double acc = 0;
    const double* v = V + size_t(q) * n;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += nb * 256) acc += v[k] * w[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[size_t(q) * nb + blockIdx.x] = red[0];
}

// partial[(j + 1) * nb + blk] = block sum of w . w
__global__ __launch_bounds__(256) void k_dot_self_j(int n, const double* __restrict__ w,
                                                    const int* __restrict__ jp,
                                                    double* __restrict__ partial) {
    __shared__ double red[256];
    const int nb = gridDim.x;
    double acc = 0;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += nb * 256) acc += w[k] * w[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[size_t(*jp + 1) * nb + blockIdx.x] = red[0];
}

// out[q] = sum of partial row q, for q <= j + extra (grid = m + 2)
__global__ __launch_bounds__(256) void k_reduce_j(const double* __restrict__ partial, int nb,
                                                  const int* __restrict__ jp, int first_extra,
                                                  double* __restrict__ out) {
    const int j = *jp;
    // first_extra = 0: rows 0..j; first_extra = 1: row j + 1 only
    const int q = first_extra ? j + 1 : blockIdx.x;
    if (first_extra ? blockIdx.x != 0 : q > j) return;
    __shared__ double red[256];
    double acc = 0;
    for (int k = threadIdx.x; k < nb; k += 256) acc += partial[size_t(q) * nb + k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

// w -= sum_{q <= j} h[q] V[q]
__global__ __launch_bounds__(256) void k_gemv_sub_j(int n, const double* __restrict__ V,
                                                    const int* __restrict__ jp,
                                                    const double* __restrict__ h,
                                                    double* __restrict__ w) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int nc = *jp + 1;
    double acc = 0;
    for (int q = 0; q < nc; ++q) acc += h[q] * V[size_t(q) * n + k];
    w[k] -= acc;
}

// V[j + 1] = w / |w| with |w| = sqrt(max(0, h[j + 1])) and a zero norm
// replaced by the smallest normal double (the NEAR0 guard of beep.cpp:1366)
__global__ void k_scale_store_j(int n, const double* __restrict__ w, const double* __restrict__ h,
                                const int* __restrict__ jp, double* __restrict__ V) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = *jp;
    const double nv = sqrt(fmax(0.0, h[j + 1]));
    const double sc = 1.0 / (nv == 0 ? 2.2250738585072014e-308 : nv);
    V[size_t(j + 1) * n + k] = w[k] * sc;
}

// h[q] += h2[q], q <= j (the second Gram-Schmidt pass's coefficients)
__global__ void k_add_j(const double* __restrict__ h2, const int* __restrict__ jp, double* __restrict__ h) {
    const int q = threadIdx.x + blockIdx.x * blockDim.x;
    if (q <= *jp) h[q] += h2[q];
}

void launch_arnoldi_step(hipStream_t s, int n, int m, int nb, double* V, const int* jp,
                         double* partial, double* h, double* w, double* h2) {
    const dim3 rows((n + 255) / 256);
    hipLaunchKernelGGL(k_multi_dot_j, dim3(nb, m + 1), dim3(256), 0, s, n, V, jp, w, partial);
    hipLaunchKernelGGL(k_reduce_j, dim3(m + 1), dim3(256), 0, s, partial, nb, jp, 0, h);
    hipLaunchKernelGGL(k_gemv_sub_j, rows, dim3(256), 0, s, n, V, jp, h, w);
    if (h2) {
        // classical Gram-Schmidt twice ("CGS2": the second pass restores the
        // orthogonality the first loses when w is nearly in span V)
        hipLaunchKernelGGL(k_multi_dot_j, dim3(nb, m + 1), dim3(256), 0, s, n, V, jp, w, partial);
        hipLaunchKernelGGL(k_reduce_j, dim3(m + 1), dim3(256), 0, s, partial, nb, jp, 0, h2);
        hipLaunchKernelGGL(k_gemv_sub_j, rows, dim3(256), 0, s, n, V, jp, h2, w);
        hipLaunchKernelGGL(k_add_j, dim3((m + 1 + 255) / 256), dim3(256), 0, s, h2, jp, h);
    }
    hipLaunchKernelGGL(k_dot_self_j, dim3(nb), dim3(256), 0, s, n, w, jp, partial);
    hipLaunchKernelGGL(k_reduce_j, dim3(1), dim3(256), 0, s, partial, nb, jp, 1, h);
    hipLaunchKernelGGL(k_scale_store_j, rows, dim3(256), 0, s, n, w, h, jp, V);
    hip_check(hipGetLastError(), "arnoldi step launch");
}

void launch_copy_col(hipStream_t s, int n, const double* V, const int* jp, double* out) {
    hipLaunchKernelGGL(k_copy_col, dim3((n + 255) / 256), dim3(256), 0, s, n, V, jp, out);
    hip_check(hipGetLastError(), "k_copy_col launch");
}

// Measurement aid: dst row r = src row idx[r], rows of 64 bytes (four float4
// per row, one per thread), so every row of src is read once in the order idx
// gives (the counters' view of a scattered read of known size).
__global__ __launch_bounds__(256) void k_gather_rows64(const float4* __restrict__ src, const int* __restrict__ idx,
                                                      size_t rows, float4* __restrict__ dst) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * 4) return;
    const size_t r = t >> 2;
    dst[t] = src[size_t(idx[r]) * 4 + (t & 3)];
}

void launch_gather_rows64(hipStream_t s, const void* src, const int* idx, size_t rows, void* dst) {
    if (rows == 0) return;
    const size_t threads = rows * 4;
    hipLaunchKernelGGL(k_gather_rows64, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s,
                       static_cast<const float4*>(src), idx, rows, static_cast<float4*>(dst));
    hip_check(hipGetLastError(), "k_gather_rows64 launch");
}

void launch_set_int(hipStream_t s, int* p, int v) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, p, v);
    hip_check(hipGetLastError(), "k_set_int launch");
}

void launch_inc_int(hipStream_t s, int* p) {
    hipLaunchKernelGGL(k_inc_int, dim3(1), dim3(64), 0, s, p);
    hip_check(hipGetLastError(), "k_inc_int launch");
}

// Node-patch point sets generated on the device from the surface
// (BasicNodePatch point lists, node_patch.cpp:55-273; sub-triangle points,
// triangle.h:387-479): patch p is vertex p, and for every triangle around
// it (in the listed order) the two sub-triangles (centre, vertex, edge
// midpoint) carry the points of the rule, weight w * (triangle area) / 6.
// A curved triangle (ctl_idx >= 0) maps the same parametric points onto its
// surface S(u, v) (curved.hpp): normal unit(dS/du x dS/dv), weight
// w * |dS/du x dS/dv| / 12, the flat triangle's w * area / 6 when S is flat.
// One thread per patch writes its run of points in the local frame (offsets
// from the patch centroid, FP64, and the FP32 copy rounded once).
__global__ __launch_bounds__(256) void k_gen_points(
    int n, const double* __restrict__ verts, const int* __restrict__ tris,
    const int* __restrict__ vt_ptr, const int* __restrict__ vt_tri,
    const int* __restrict__ vt_corner, const int* __restrict__ ctl_idx,
    const double* __restrict__ ctl, const Patch64* __restrict__ patches,
    const int* __restrict__ ptr, const RuleDev rule, int per_area, Pt64* __restrict__ out64,
    Pt32* __restrict__ out32) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Patch64 pa = patches[p];
    int k = ptr[p];
    if (rule.k == 0) {
        const double w = per_area ? 1.0 : pa.area;
        if (out64) out64[k] = Pt64{0.0, 0.0, 0.0, w, pa.nx, pa.ny, pa.nz, 0.0};
        if (out32)
            out32[k] = Pt32{0.f, 0.f, 0.f, float(w), float(pa.nx), float(pa.ny), float(pa.nz), 0.f};
        return;
    }
    const double wscale = per_area ? 1.0 / pa.area : 1.0;
    for (int t = vt_ptr[p]; t < vt_ptr[p + 1]; ++t) {
        const int tri = vt_tri[t], corner = vt_corner[t];
        const int i1 = tris[3 * tri], i2 = tris[3 * tri + 1], i3 = tris[3 * tri + 2];
        const double x1 = verts[3 * i1], y1 = verts[3 * i1 + 1], z1 = verts[3 * i1 + 2];
        const double x2 = verts[3 * i2], y2 = verts[3 * i2 + 1], z2 = verts[3 * i2 + 2];
        const double x3 = verts[3 * i3], y3 = verts[3 * i3 + 1], z3 = verts[3 * i3 + 2];
        const double ax = x2 - x1, ay = y2 - y1, az = z2 - z1;
        const double bx = x3 - x1, by = y3 - y1, bz = z3 - z1;
        const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const double cn = sqrt(cx * cx + cy * cy + cz * cz);
        const double tri_area = 0.5 * cn;
        const double nx = cx / cn, ny = cy / cn, nz = cz / cn;
        // barycentric (u, v, w) of the vertex, of the midpoints of its two edges
        // and of the triangle centre; a point is w*v1 + u*v2 + v*v3
        double o[3], m1[3], m2[3];
        const double third = 1.0 / 3.0;
        if (corner == 0) {
            o[0] = 0; o[1] = 0; o[2] = 1; m1[0] = 0.5; m1[1] = 0; m1[2] = 0.5; m2[0] = 0; m2[1] = 0.5; m2[2] = 0.5;
        } else if (corner == 1) {
            o[0] = 1; o[1] = 0; o[2] = 0; m1[0] = 0.5; m1[1] = 0.5; m1[2] = 0; m2[0] = 0.5; m2[1] = 0; m2[2] = 0.5;
        } else {
            o[0] = 0; o[1] = 1; o[2] = 0; m1[0] = 0; m1[1] = 0.5; m1[2] = 0.5; m2[0] = 0.5; m2[1] = 0.5; m2[2] = 0;
        }
        const double c3[3] = {third, third, third};
        const int ci = ctl_idx ? ctl_idx[tri] : -1;
        for (int s = 0; s < 2; ++s) {
            const double* P1 = s == 0 ? c3 : m2;
            const double* P3 = s == 0 ? m1 : c3;
            for (int j = 0; j < rule.k; ++j) {
                const double ru = rule.u[j], rv = rule.v[j], r0 = 1.0 - ru - rv;
                const double bu = P1[0] * r0 + o[0] * ru + P3[0] * rv;
                const double bv = P1[1] * r0 + o[1] * ru + P3[1] * rv;
                const double bw = P1[2] * r0 + o[2] * ru + P3[2] * rv;
                double px, py, pz, wt, qx = nx, qy = ny, qz = nz;
                if (ci < 0) {
                    px = x1 * bw + x2 * bu + x3 * bv - pa.cx;
                    py = y1 * bw + y2 * bu + y3 * bv - pa.cy;
                    pz = z1 * bw + z2 * bu + z3 * bv - pa.cz;
                    wt = rule.w[j] * tri_area / 6.0 * wscale;
                } else {
                    double q[3], du[3], dv[3];
                    beepcurve::eval(ctl + 63 * size_t(ci), bu, bv, q, du, dv);
                    px = q[0] - pa.cx;
                    py = q[1] - pa.cy;
                    pz = q[2] - pa.cz;
                    const double ex = du[1] * dv[2] - du[2] * dv[1];
                    const double ey = du[2] * dv[0] - du[0] * dv[2];
                    const double ez = du[0] * dv[1] - du[1] * dv[0];
                    const double en = sqrt(ex * ex + ey * ey + ez * ez);
                    qx = ex / en;
                    qy = ey / en;
                    qz = ez / en;
                    const double uu = du[0] * du[0] + du[1] * du[1] + du[2] * du[2];
                    const double vv = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
                    const double uv = du[0] * dv[0] + du[1] * dv[1] + du[2] * dv[2];
                    const double jac = sqrt(uu * vv - uv * uv);
                    wt = rule.w[j] / 6.0 * jac / 2.0;
                    if (per_area) wt /= pa.area;
                }
                if (out64) out64[k] = Pt64{px, py, pz, wt, qx, qy, qz, 0.0};
                if (out32)
                    out32[k] = Pt32{float(px), float(py), float(pz), float(wt), float(qx), float(qy),
                                    float(qz), 0.f};
                ++k;
            }
        }
    }
}

void launch_gen_points(hipStream_t s, int n, const double* verts, const int* tris,
                       const int* vt_ptr, const int* vt_tri, const int* vt_corner,
                       const int* ctl_idx, const double* ctl, const Patch64* patches,
                       const int* ptr, const RuleDev& rule, int per_area, Pt64* out64,
                       Pt32* out32) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gen_points, dim3((n + 255) / 256), dim3(256), 0, s, n, verts, tris,
                       vt_ptr, vt_tri, vt_corner, ctl_idx, ctl, patches, ptr, rule, per_area,
                       out64, out32);
    hip_check(hipGetLastError(), "k_gen_points launch");
}

// Points k0 .. k0+n-1 of a CSR point set (raw: 7 doubles x, y, z, nx, ny,
// nz, w each) into the local frame of the patch that owns them (offsets from
// its centroid, formed in FP64 and rounded once for the FP32 copy).
__global__ void k_to_local(int n, int k0, const double* __restrict__ raw,
                           const int* __restrict__ ptr, int N,
                           const Patch64* __restrict__ patches, Pt64* __restrict__ out64,
                           Pt32* __restrict__ out32) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = k0 + t;
    int lo = 0, hi = N;  // patch i with ptr[i] <= k < ptr[i + 1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ptr[mid] <= k) lo = mid;
        else hi = mid;
    }
    const Patch64 p = patches[lo];
    const double* d = raw + 7 * size_t(t);
    const double x = d[0] - p.cx, y = d[1] - p.cy, z = d[2] - p.cz;
    if (out64) out64[k] = Pt64{x, y, z, d[6], d[3], d[4], d[5], 0.0};
    if (out32)
        out32[k] = Pt32{float(x), float(y), float(z), float(d[6]), float(d[3]), float(d[4]),
                        float(d[5]), 0.0f};
}

void launch_to_local(hipStream_t s, int n, int k0, const double* raw, const int* ptr, int N,
                     const Patch64* patches, Pt64* out64, Pt32* out32) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_local, dim3((n + 255) / 256), dim3(256), 0, s, n, k0, raw, ptr, N,
                       patches, out64, out32);
    hip_check(hipGetLastError(), "k_to_local launch");
}

}  // namespace beepdev

namespace beepdev {

// Self-test of the 16x16x4 MFMA operand layout used by the M2L GEMM:
// A (16x4) lane l holds A[l % 16][l / 16], B (4x16) lane l holds
// B[l / 16][l % 16]; out[4 l + r] = accumulator element r of lane l for
// D = A B, A[i][k] = a[i*4+k], B[k][j] = b[k*16+j].
typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_mfma_selftest(const double* __restrict__ a,
                                                      const double* __restrict__ b, int fp32,
                                                      double* __restrict__ out) {
    const int l = threadIdx.x;
    const int i = l % 16, k = l / 16;
    if (fp32) {
        f4v acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(float(a[i * 4 + k]), float(b[k * 16 + i]), acc,
                                                   0, 0, 0);
        for (int r = 0; r < 4; ++r) out[4 * l + r] = double(acc[r]);
    } else {
        d4v acc = {0.0, 0.0, 0.0, 0.0};
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i * 4 + k], b[k * 16 + i], acc, 0, 0, 0);
        for (int r = 0; r < 4; ++r) out[4 * l + r] = acc[r];
    }
}

void launch_mfma_selftest(hipStream_t s, const double* a, const double* b, int fp32, double* out) {
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, s, a, b, fp32, out);
    hip_check(hipGetLastError(), "k_mfma_selftest launch");
}

}  // namespace beepdev
