// Green's-function kernels (FP64):
//  * a_adag as a gather over the target sector (replaces the signed scatter of
//    SpinResolvedStorage::a_adag, SpinResolvedStorage.h:293-381): creating or
//    destroying one electron changes only one spin's pattern, so the source
//    index and sign factor into an up part and a down part;
//  * ChiLoc::operation (ChiLoc.h:474-530): the operator is diagonal and its
//    value is a sum of an up-pattern part and a down-pattern part;
//  * Lanczos::get_frac_point (Lanczos.h:201-224): one thread per energy point.
#include <cfloat>

#include "edh_internal.h"
#include "edh_kernels.h"

namespace edh {

// out = c(+)|in> (accumulate: out = c(+)|in> + out, the order in which the
// reference sums the vectors of several orbitals, GreensFunction.h:434-438)
__global__ void separable_gather_kernel(const double* in, int64_t ldin, const int32_t* mu, const int32_t* md,
                                        const int8_t* su, const int8_t* sd, double* out, int64_t Dup, int64_t Ddn,
                                        int accumulate) {
  const int64_t n = Dup * Ddn;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = e / Ddn, d = e - u * Ddn;
    const int32_t a = mu[u], b = md[d];
    double x = 0.0;
    if (a >= 0 && b >= 0) x = (double)(su[u] * sd[d]) * in[(int64_t)a * ldin + b];
    out[e] = accumulate ? x + out[e] : x;
  }
}

__global__ void separable_scale_kernel(const double* in, const double* fu, const double* fd, double* out, int64_t Dup,
                                       int64_t Ddn) {
  const int64_t n = Dup * Ddn;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = e / Ddn, d = e - u * Ddn;
    out[e] = (fu[u] + fd[d]) * in[e];
  }
}

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
// Smith's algorithm for a / b.
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, den = b.re + b.im * r;
    return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  }
  const double r = b.re / b.im, den = b.re * r + b.im;
  return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}

__global__ void cf_kernel(int nz, const double* zre, const double* zim, double ev, int nlanc, int isign, double expb,
                          const double* alpha, const double* beta, double* out_re, double* out_im) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nz) return;
  const cplx z = {zre[q], zim[q]};
  const double big = (double)FLT_MAX / 2.0;
  const double shift = 1.0 / ((double)FLT_MAX / 1000.0);
  cplx res;
  if (nlanc == 1) {
    const cplx dl0 = {z.re - alpha[0] * isign, z.im};
    res = cdiv({ev * expb, 0.0}, dl0);
  } else {
    const int n = nlanc;
    const cplx dl1 = {z.re - alpha[n - 1] * isign, z.im};
    const cplx dl2 = {z.re - alpha[n - 2] * isign, z.im};
    cplx dA = dl1;                                                    // det[i+2]
    cplx dB = csub(cmul(dl2, dl1), {beta[n - 1] * beta[n - 1], 0.0});  // det[i+1]
    for (int i = n - 3; i >= 0; --i) {
      const cplx dli = {z.re - alpha[i] * isign, z.im};
      const double b2 = beta[i + 1] * beta[i + 1];
      cplx d = csub(cmul(dli, dB), cscale(dA, b2));
      if (hypot(d.re, d.im) > big && i != 0) {
        d = cscale(d, shift);
        dB = cscale(dB, shift);
      }
      dA = dB;
      dB = d;
    }
    res = cdiv(cscale(dA, ev * expb), dB);
  }
  out_re[q] = res.re;
  out_im[q] = res.im;
}

// Many continued fractions at once: block row j evaluates job j at the
// points z = excited_j * isign_j + i wn[q] (the same recurrence as cf_kernel).
__global__ void cf_batch_kernel(int nz, const double* wr, const double* wn, const CfJob* jobs, double* out_re,
                                double* out_im) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const CfJob& J = jobs[blockIdx.y];
  if (q >= nz) return;
  // z = freq_point + isign * E (Lanczos.h:146-148): i w_n on the Matsubara
  // mesh (wr null), w + i pi/beta on the real axis
  const cplx z = {(wr ? wr[q] : 0.0) + J.excited * J.isign, wn[q]};
  const double big = (double)FLT_MAX / 2.0;
  const double shift = 1.0 / ((double)FLT_MAX / 1000.0);
  const double* alpha = J.alpha;
  const double* beta = J.beta;
  const int isign = J.isign;
  cplx res;
  if (J.nlanc == 1) {
    const cplx dl0 = {z.re - alpha[0] * isign, z.im};
    res = cdiv({J.ev * J.expb, 0.0}, dl0);
  } else {
    const int n = J.nlanc;
    const cplx dl1 = {z.re - alpha[n - 1] * isign, z.im};
    const cplx dl2 = {z.re - alpha[n - 2] * isign, z.im};
    cplx dA = dl1;
    cplx dB = csub(cmul(dl2, dl1), {beta[n - 1] * beta[n - 1], 0.0});
    for (int i = n - 3; i >= 0; --i) {
      const cplx dli = {z.re - alpha[i] * isign, z.im};
      const double b2 = beta[i + 1] * beta[i + 1];
      cplx d = csub(cmul(dli, dB), cscale(dA, b2));
      if (hypot(d.re, d.im) > big && i != 0) {
        d = cscale(d, shift);
        dB = cscale(dB, shift);
      }
      dA = dB;
      dB = d;
    }
    res = cdiv(cscale(dA, J.ev * J.expb), dB);
  }
  out_re[(size_t)blockIdx.y * nz + q] = res.re;
  out_im[(size_t)blockIdx.y * nz + q] = res.im;
}

void dev_cf_batch(int njobs, int nz, const double* wr, const double* wn, const CfJob* jobs, double* out_re,
                  double* out_im, hipStream_t s) {
  if (njobs <= 0 || nz <= 0) return;
  hipLaunchKernelGGL(cf_batch_kernel, dim3((nz + 127) / 128, njobs), dim3(128), 0, s, nz, wr, wn, jobs, out_re,
                     out_im);
  EDH_HIP(hipGetLastError());
}

void dev_separable_gather(const double* in, int64_t ldin, const int32_t* mu, const int32_t* md, const int8_t* su,
                          const int8_t* sd, double* out, int64_t Dup, int64_t Ddn, hipStream_t s, bool accumulate) {
  if (Dup * Ddn == 0) return;
  hipLaunchKernelGGL(separable_gather_kernel, dim3(reduce_grid(Dup * Ddn)), dim3(kVecBlock), 0, s, in, ldin, mu, md,
                     su, sd, out, Dup, Ddn, accumulate ? 1 : 0);
  EDH_HIP(hipGetLastError());
}

void dev_separable_scale(const double* in, const double* fu, const double* fd, double* out, int64_t Dup, int64_t Ddn,
                         hipStream_t s) {
  if (Dup * Ddn == 0) return;
  hipLaunchKernelGGL(separable_scale_kernel, dim3(reduce_grid(Dup * Ddn)), dim3(kVecBlock), 0, s, in, fu, fd, out,
                     Dup, Ddn);
  EDH_HIP(hipGetLastError());
}

void dev_cf_eval(int nz, const double* zre, const double* zim, double ev, int nlanc, int isign, double expb,
                 const double* alpha, const double* beta, double* out_re, double* out_im, hipStream_t s) {
  if (nz <= 0) return;
  hipLaunchKernelGGL(cf_kernel, dim3((nz + 255) / 256), dim3(256), 0, s, nz, zre, zim, ev, nlanc, isign, expb, alpha,
                     beta, out_re, out_im);
  EDH_HIP(hipGetLastError());
}

}  // namespace edh
