// coupled2d.hip -- medium of the coupled P-S potential equations, built on the GPU at plan setup.
//
// init_coupled_medium (src/equations/coupled2d/medium.jl:50-117) on the padded grid: alpha = vp^2
// and beta = vs^2 enter the propagation terms raw; the coupling terms use the first derivatives
// and Laplacians of Gaussian-smoothed copies.  The reference builds these on the host; here they
// are four setup launches over the padded grid (square, smoothing along x, smoothing along z,
// derivatives), each cell's sum taken term by term in the reference's order, so the Float32 results
// are the same bits as a sequential host loop (the library is compiled with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fomo_common.h"
#include "plan.h"

namespace fomo {

namespace {

constexpr int kBlock = 256;

struct DerivW {
  float d1[kMaxN], d2[kMaxN];
  float c0;
  int N;
};

__global__ void square_kernel(const float* __restrict__ v, float* __restrict__ out, size_t n) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
    const float x = v[c];
    out[c] = x * x;
  }
}

// One pass of the separable Gaussian (medium.jl:132-151): out(i,j) = sum_k ker[k] f(clamp(i+k), j)
// along x (dir 0) or f(i, clamp(j+k)) along z (dir 1), k = -R..R, accumulated from zero.
__global__ void smooth_kernel(const float* __restrict__ f, float* __restrict__ out, int nx, int nz,
                              const float* __restrict__ ker, int R, int dir) {
  const size_t n = (size_t)nx * nz;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(c / (size_t)nx), i = (int)(c - (size_t)j * nx);
    float v = 0.0f;
    for (int k = -R; k <= R; ++k) {
      const size_t s = dir == 0 ? (size_t)min(max(i + k, 0), nx - 1) + (size_t)j * nx
                                : (size_t)i + (size_t)min(max(j + k, 0), nz - 1) * nx;
      v += ker[k + R] * f[s];
    }
    out[c] = v;
  }
}

// medium.jl:156-193: d/dx on N <= i < nx-N (every j), d/dz on N <= j < nz-N (every i), each
// divided by h; the Laplacian on both ranges, scaled by 1/(h h); zero elsewhere.
__global__ void derivs_kernel(const float* __restrict__ f, float* __restrict__ gx, float* __restrict__ gz,
                              float* __restrict__ lap, int nx, int nz, const DerivW W, float h, float ih2) {
  const size_t n = (size_t)nx * nz;
  const int N = W.N;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(c / (size_t)nx), i = (int)(c - (size_t)j * nx);
    const bool ix = i >= N && i < nx - N, iz = j >= N && j < nz - N;
    float vx = 0.0f, vz = 0.0f, vl = 0.0f;
    if (ix) {
      float v = 0.0f;
      for (int l = 1; l <= N; ++l) v += W.d1[l - 1] * (f[c + l] - f[c - l]);
      vx = v / h;
    }
    if (iz) {
      float v = 0.0f;
      for (int l = 1; l <= N; ++l) v += W.d1[l - 1] * (f[c + (size_t)l * nx] - f[c - (size_t)l * nx]);
      vz = v / h;
    }
    if (ix && iz) {
      float ax = W.c0 * f[c], az = W.c0 * f[c];
      for (int l = 1; l <= N; ++l) ax += W.d2[l - 1] * (f[c + l] + f[c - l]);
      for (int l = 1; l <= N; ++l) az += W.d2[l - 1] * (f[c + (size_t)l * nx] + f[c - (size_t)l * nx]);
      vl = (ax + az) * ih2;
    }
    gx[c] = vx;
    gz[c] = vz;
    lap[c] = vl;
  }
}

int grid_for(size_t n) { return (int)std::min<size_t>(8192, (n + kBlock - 1) / kBlock); }

// Device scratch released when the builder returns.
struct Scratch {
  std::vector<void*> p;
  ~Scratch() {
    for (void* q : p) (void)hipFree(q);
  }
  int get(float** out, size_t n) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, (n ? n : 1) * sizeof(float));
    if (e != hipSuccess)
      return set_error(e == hipErrorOutOfMemory ? FOMO_ENOMEM : FOMO_EHIP,
                       std::string("coupled medium scratch: ") + hipGetErrorString(e));
    p.push_back(v);
    *out = (float*)v;
    return FOMO_OK;
  }
};

}  // namespace

// Gaussian taps of medium.jl:118-131: radius ceil(3 sigma), exp(-0.5 (k/sigma)^2) in Float64
// rounded to Float32, divided by their Float32 sum taken first to last (the order of Julia's
// sequential reduction).
std::vector<float> coupled_gauss_taps(double sigma) {
  const int R = (int)std::ceil(3.0 * sigma);
  std::vector<float> ker(2 * R + 1);
  for (int k = -R; k <= R; ++k) {
    const double t = (double)k / sigma;
    ker[k + R] = (float)std::exp(-0.5 * (t * t));
  }
  float s = ker[0];
  if (ker.size() > 1) {
    s = ker[0] + ker[1];
    for (size_t k = 2; k < ker.size(); ++k) s += ker[k];
  }
  for (float& v : ker) v = v / s;
  return ker;
}

// The eight medium arrays of a padded (nx, nz) grid from the padded vp, vs (host), as new plan
// allocations in `out`: alpha, beta, d alpha/dx, d alpha/dz, d beta/dx, d beta/dz, lap alpha,
// lap beta.  sigma <= 0: no smoothing.
int coupled_medium2d(PlanBase* p, const std::vector<float>& vpp, const std::vector<float>& vsp, int nx, int nz,
                     int fd_order, float dh, double sigma, float* out[8]) {
  const size_t n = (size_t)nx * nz;
  if (sigma > 0.0 && !(sigma <= 1.0e4))
    return set_error(FOMO_EINVAL, "smooth_sigma must be finite and at most 1e4 grid points");
  DerivW W{};
  W.N = fd_order / 2;
  if (fomo_centered_d1(fd_order, W.d1) || fomo_centered_d2(fd_order, &W.c0, W.d2))
    return set_error(FOMO_EINVAL, "Unsupported FD order");
  int rc;
  for (int m = 0; m < 8; ++m)
    if ((rc = p->alloc(&out[m], n))) return rc;
  Scratch sc;
  float *vp_d = nullptr, *vs_d = nullptr, *as = nullptr, *bs = nullptr, *tmp = nullptr, *ker_d = nullptr;
  if ((rc = sc.get(&vp_d, n)) || (rc = sc.get(&vs_d, n))) return rc;
  FOMO_HIP_TRY(hipMemcpyAsync(vp_d, vpp.data(), n * sizeof(float), hipMemcpyHostToDevice, p->stream));
  FOMO_HIP_TRY(hipMemcpyAsync(vs_d, vsp.data(), n * sizeof(float), hipMemcpyHostToDevice, p->stream));
  const int g = grid_for(n);
  hipLaunchKernelGGL(square_kernel, dim3(g), dim3(kBlock), 0, p->stream, vp_d, out[0], n);
  hipLaunchKernelGGL(square_kernel, dim3(g), dim3(kBlock), 0, p->stream, vs_d, out[1], n);
  FOMO_HIP_TRY(hipGetLastError());
  const float* sa = out[0];
  const float* sb = out[1];
  std::vector<float> ker;
  if (sigma > 0.0) {
    ker = coupled_gauss_taps(sigma);
    const int R = (int)(ker.size() / 2);
    if ((rc = sc.get(&as, n)) || (rc = sc.get(&bs, n)) || (rc = sc.get(&tmp, n)) || (rc = sc.get(&ker_d, ker.size())))
      return rc;
    FOMO_HIP_TRY(hipMemcpyAsync(ker_d, ker.data(), ker.size() * sizeof(float), hipMemcpyHostToDevice, p->stream));
    hipLaunchKernelGGL(smooth_kernel, dim3(g), dim3(kBlock), 0, p->stream, out[0], tmp, nx, nz, ker_d, R, 0);
    hipLaunchKernelGGL(smooth_kernel, dim3(g), dim3(kBlock), 0, p->stream, tmp, as, nx, nz, ker_d, R, 1);
    hipLaunchKernelGGL(smooth_kernel, dim3(g), dim3(kBlock), 0, p->stream, out[1], tmp, nx, nz, ker_d, R, 0);
    hipLaunchKernelGGL(smooth_kernel, dim3(g), dim3(kBlock), 0, p->stream, tmp, bs, nx, nz, ker_d, R, 1);
    FOMO_HIP_TRY(hipGetLastError());
    sa = as;
    sb = bs;
  }
  const float h = dh, ih2 = 1.0f / (h * h);
  hipLaunchKernelGGL(derivs_kernel, dim3(g), dim3(kBlock), 0, p->stream, sa, out[2], out[3], out[6], nx, nz, W, h, ih2);
  hipLaunchKernelGGL(derivs_kernel, dim3(g), dim3(kBlock), 0, p->stream, sb, out[4], out[5], out[7], nx, nz, W, h, ih2);
  FOMO_HIP_TRY(hipGetLastError());
  // the scratch buffers are freed on return: wait for the launches that read them
  FOMO_HIP_TRY(hipStreamSynchronize(p->stream));
  return FOMO_OK;
}

}  // namespace fomo
