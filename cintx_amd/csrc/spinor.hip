// Cartesian -> spinor transform of two-electron blocks on the device.
//
// int2e_spinor (libcint-master/src/cint2e.c:1215-1224 with c2s_sf_2e1 /
// c2s_sf_2e2, cart2sph.c:5385-5528) and int3c2e_spinor (cint3c2e.c:719-727
// with c2s_sf_3c2e1, cart2sph.c:5997-6050): the spin-free two-component
// transform of the Cartesian integrals (ij|kl),
//
//   out[i][j][k][l] = sum_{abcd} sum_{s,t} conj(C_i[s][a]) C_j[s][b]
//                                          conj(C_k[t][c]) C_l[t][d] (ab|cd)
//
// (bra spinors conjugated, the spin summed per electron); for int3c2e the
// third index is the real spherical transform of k.  One workgroup per job
// (one contracted block), four passes through a per-workgroup workspace:
//   Y[s][i][b][cd] = sum_a conj(C_i[s][a]) X[a][b][cd]
//   Z[i][j][cd]    = sum_s sum_b C_j[s][b] Y[s][i][b][cd]
//   W[t][i][j][k][d] = sum_c conj(C_k[t][c]) Z[i][j][c + nfk d]
//   out[i][j][k][l]  = sum_t sum_d C_l[t][d] W[t][i][j][k][d]
// The blocks are small (at most 15^4 Cartesian elements) and L2-resident;
// the ERI kernels that produce them dominate the time.
#include "common.h"

namespace cx {

__global__ __launch_bounds__(256) void spinor_kernel(
    const SpinorJob* __restrict__ jobs, int64_t njobs, const double* __restrict__ cart,
    double2* __restrict__ out, const double2* __restrict__ tab, const double* __restrict__ c2s,
    double2* __restrict__ scratch, int64_t slice, int four) {
  double2* const A = scratch + (int64_t)blockIdx.x * slice;
  for (int64_t jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
    const SpinorJob J = jobs[jb];
    const int li = J.l[0], lj = J.l[1], lk = J.l[2], ll = J.l[3];
    const int fi = ncart(li), fj = ncart(lj), fk = ncart(lk), fl = four ? ncart(ll) : 1;
    const int di = spinor_len(li, J.kappa[0]), dj = spinor_len(lj, J.kappa[1]);
    const int dk = four ? spinor_len(lk, J.kappa[2]) : nsph(lk);
    const int dl = four ? spinor_len(ll, J.kappa[3]) : 1;
    const double2* Ci = tab + spinor_tab_off(li) + spinor_row0(li, J.kappa[0]) * 2 * fi;
    const double2* Cj = tab + spinor_tab_off(lj) + spinor_row0(lj, J.kappa[1]) * 2 * fj;
    const double* X = cart + J.src;
    const int kl = fk * fl;
    const int n1 = 2 * di * fj * kl, n2 = di * dj * kl, n3 = four ? 2 * di * dj * dk * fl : 0;
    double2* const Y = A;                       // pass 1 out, pass 3 out (W)
    double2* const Z = A + (n1 > n3 ? n1 : n3);  // pass 2 out

    // pass 1: bra i (conjugated), both spin components
    for (int e = threadIdx.x; e < n1; e += blockDim.x) {
      const int cd = e % kl;
      int r = e / kl;
      const int b = r % fj;
      r /= fj;
      const int ii = r % di, s = r / di;
      const double* x = X + (int64_t)b * J.cs[1] + (int64_t)(cd % fk) * J.cs[2] +
                        (int64_t)(cd / fk) * J.cs[3];
      const double2* co = Ci + (ii * 2 + s) * fi;
      double re = 0.0, im = 0.0;
      for (int a = 0; a < fi; ++a) {
        const double v = x[(int64_t)a * J.cs[0]];
        re += co[a].x * v;
        im -= co[a].y * v;
      }
      Y[e] = make_double2(re, im);
    }
    __syncthreads();
    // pass 2: ket j, spin summed
    for (int e = threadIdx.x; e < n2; e += blockDim.x) {
      const int cd = e % kl;
      const int r = e / kl;
      const int jj = r % dj, ii = r / dj;
      double re = 0.0, im = 0.0;
      for (int s = 0; s < 2; ++s) {
        const double2* co = Cj + (jj * 2 + s) * fj;
        const double2* y = Y + ((int64_t)(s * di + ii) * fj) * kl + cd;
        for (int b = 0; b < fj; ++b) {
          const double2 c = co[b], v = y[(int64_t)b * kl];
          re += c.x * v.x - c.y * v.y;
          im += c.x * v.y + c.y * v.x;
        }
      }
      Z[e] = make_double2(re, im);
    }
    __syncthreads();
    if (four) {
      const double2* Ck = tab + spinor_tab_off(lk) + spinor_row0(lk, J.kappa[2]) * 2 * fk;
      const double2* Cl = tab + spinor_tab_off(ll) + spinor_row0(ll, J.kappa[3]) * 2 * fl;
      // pass 3: bra k (conjugated), both spin components
      for (int e = threadIdx.x; e < n3; e += blockDim.x) {
        const int d = e % fl;
        int r = e / fl;
        const int kk = r % dk;
        r /= dk;
        const int ij = r % (di * dj), t = r / (di * dj);
        const double2* co = Ck + (kk * 2 + t) * fk;
        const double2* z = Z + (int64_t)ij * kl + (int64_t)fk * d;
        double re = 0.0, im = 0.0;
        for (int c = 0; c < fk; ++c) {
          const double2 q = co[c], v = z[c];
          re += q.x * v.x + q.y * v.y;
          im += q.x * v.y - q.y * v.x;
        }
        Y[e] = make_double2(re, im);
      }
      __syncthreads();
      // pass 4: ket l, spin summed, into the caller's layout
      const int n4 = di * dj * dk * dl;
      for (int e = threadIdx.x; e < n4; e += blockDim.x) {
        const int ii = e % di;
        int r = e / di;
        const int jj = r % dj;
        r /= dj;
        const int kk = r % dk, lq = r / dk;
        const int ijk = (ii * dj + jj) * dk + kk;
        double re = 0.0, im = 0.0;
        for (int t = 0; t < 2; ++t) {
          const double2* co = Cl + (lq * 2 + t) * fl;
          const double2* w = Y + ((int64_t)t * di * dj * dk + ijk) * fl;
          for (int d = 0; d < fl; ++d) {
            const double2 c = co[d], v = w[d];
            re += c.x * v.x - c.y * v.y;
            im += c.x * v.y + c.y * v.x;
          }
        }
        out[J.dst + ii + (int64_t)jj * J.so[0] + (int64_t)kk * J.so[1] + (int64_t)lq * J.so[2]] =
            make_double2(re, im);
      }
    } else {
      // k: real spherical transform (rows m of the c2s table)
      const double* R = c2s + (int64_t)lk * kC2sM * kC2sC;
      const int n4 = di * dj * dk;
      for (int e = threadIdx.x; e < n4; e += blockDim.x) {
        const int ii = e % di;
        const int r = e / di;
        const int jj = r % dj, kk = r / dj;
        const double* rk = R + kk * kC2sC;
        const double2* z = Z + (int64_t)(ii * dj + jj) * kl;
        double re = 0.0, im = 0.0;
        for (int c = 0; c < fk; ++c) {
          re += rk[c] * z[c].x;
          im += rk[c] * z[c].y;
        }
        out[J.dst + ii + (int64_t)jj * J.so[0] + (int64_t)kk * J.so[1]] = make_double2(re, im);
      }
    }
    __syncthreads();  // the next job reuses the workspace
  }
}

hipError_t launch_spinor(const SpinorJob* jobs, int64_t njobs, const double* cart, double* out,
                         const double* tab, const double* c2s, double* scratch, int64_t slice,
                         int grid, bool four, hipStream_t s) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(spinor_kernel, dim3((unsigned)grid), dim3(256), 0, s, jobs, njobs, cart,
                     (double2*)out, (const double2*)tab, c2s, (double2*)scratch, slice,
                     four ? 1 : 0);
  return hipGetLastError();
}

}  // namespace cx
