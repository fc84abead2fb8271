// Determinant head of the Psiformer with forward-mode jets, float64 throughout.
//
// Per configuration (one workgroup):
//   env[a,i]   = sum_G W[a,G,s_i] exp(i G.r_i)                      (psiformer.py:145-148)
//   M_k[a,i]   = env[a,i] (1 + bf[k,a,i])                           (:151)
//   logdet_k   = sum log|u_pp| + log(prod u_pp/|u_pp| * parity)      (:152-153, LU with partial
//                pivoting on |re|+|im| as LAPACK getrf does)
//   tangents   d_c logdet = tr(M^-1 dM_c)
//   Laplacian  Lap logdet = tr(M^-1 Lap M) - sum_c tr((M^-1 dM_c)^2)
//   log psi    = logsumexp_k(logdet_k + w_k) (K > 1, :155-158) + Jastrow (:162-188)
// dM_c = denv_c (1 + bf) + env dbf_c, where denv_c is non-zero only in the column of the
// electron that tangent c moves; Lap M = Lap env (1+bf) + env Lap bf + 2 sum_c denv_c dbf_c.
#include "nqs_common.h"
#include "kernels.h"
#include "bwd.h"
#include "tune.h"

namespace nqs {

__device__ inline void block_sum2(double& a, double& b, double* red) {
  // sum (a, b) over the workgroup; every thread receives the result.  red: >= 2*8 doubles
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  __syncthreads();
  if (lane == 0) {
    red[2 * wave] = a;
    red[2 * wave + 1] = b;
  }
  __syncthreads();
  a = 0;
  b = 0;
  for (int w = 0; w < nw; ++w) {
    a += red[2 * w];
    b += red[2 * w + 1];
  }
  __syncthreads();
}


// LU with partial pivoting of the N x N complex matrix M (LDS, row-major) by the whole
// workgroup; pivot = largest |re|+|im| in the column (LAPACK getrf's choice).  piv[q] = row
// swapped with q; flag[0] = parity of the swaps, flag[1] = 1 if a pivot was exactly zero.
__device__ void wg_lu(cd* M, int N, int* piv, int* flag) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (tid == 0) {
    flag[0] = 0;
    flag[1] = 0;
  }
  __syncthreads();
  for (int q = 0; q < N; ++q) {
    if (tid == 0) {
      int best = q;
      double bv = fabs(M[q * N + q].re) + fabs(M[q * N + q].im);
      for (int r = q + 1; r < N; ++r) {
        const double v = fabs(M[r * N + q].re) + fabs(M[r * N + q].im);
        if (v > bv) {
          bv = v;
          best = r;
        }
      }
      piv[q] = best;
      if (best != q) flag[0] ^= 1;
      if (bv == 0.0) flag[1] = 1;
    }
    __syncthreads();
    const int pr = piv[q];
    if (pr != q)
      for (int c = tid; c < N; c += nthr) {
        const cd t = M[q * N + c];
        M[q * N + c] = M[pr * N + c];
        M[pr * N + c] = t;
      }
    __syncthreads();
    const cd pv = M[q * N + q];
    const bool zero = (pv.re == 0.0 && pv.im == 0.0);
    for (int r = q + 1 + tid; r < N; r += nthr) M[r * N + q] = zero ? mk(0, 0) : cdiv_(M[r * N + q], pv);
    __syncthreads();
    const int w_ = N - q - 1;
    for (int e = tid; e < w_ * w_; e += nthr) {
      const int r = q + 1 + e / w_, c = q + 1 + e % w_;
      M[r * N + c] = M[r * N + c] - M[r * N + q] * M[q * N + c];
    }
    __syncthreads();
  }
}

// Mi = U^-1 L^-1 P from the LU factors in M; column c computed by thread c.  Ends synced.
__device__ void wg_inverse(const cd* M, cd* Mi, int N, const int* piv) {
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    for (int r = 0; r < N; ++r) Mi[r * N + c] = mk(r == c ? 1.0 : 0.0, 0);
    for (int q = 0; q < N; ++q) {
      const int pr = piv[q];
      if (pr != q) {
        const cd t = Mi[q * N + c];
        Mi[q * N + c] = Mi[pr * N + c];
        Mi[pr * N + c] = t;
      }
    }
    for (int r = 1; r < N; ++r) {
      cd s = Mi[r * N + c];
      for (int q = 0; q < r; ++q) s = s - M[r * N + q] * Mi[q * N + c];
      Mi[r * N + c] = s;
    }
    for (int r = N - 1; r >= 0; --r) {
      cd s = Mi[r * N + c];
      for (int q = r + 1; q < N; ++q) s = s - M[r * N + q] * Mi[q * N + c];
      Mi[r * N + c] = cdiv_(s, M[r * N + r]);
    }
  }
  __syncthreads();
}

// log det from LU factors: sum log|u_qq| + log(prod u_qq/|u_qq| * parity) (principal branch).
__device__ inline cd lu_logdet(const cd* M, int N, int parity) {
  double la = 0, sre = parity ? -1.0 : 1.0, sim = 0;
  for (int q = 0; q < N; ++q) {
    const cd u = M[q * N + q];
    const double au = cabs_(u);
    la += log(au);
    const double ur = au > 0 ? u.re / au : 0.0, ui = au > 0 ? u.im / au : 0.0;
    const double nr = sre * ur - sim * ui, ni = sre * ui + sim * ur;
    sre = nr;
    sim = ni;
  }
  return mk(la + log(hypot(sre, sim)), atan2(sim, sre));
}

// Jastrow pair function u(r) and its first two radial derivatives.
__device__ inline void jastrow_u(const DetParams& p, bool same, double r, double& u, double& u1, double& u2) {
  u = u1 = u2 = 0;
  if (p.jastrow_kind == 0) {
    const double* c = p.jas + (same ? 0 : p.n_knots);
    const double w = p.rbf_w;
    for (int k = 0; k < p.n_knots; ++k) {
      const double mu = (p.n_knots > 1) ? p.rbf_rcut * (double)k / (double)(p.n_knots - 1) : 0.0;
      const double z = (r - mu) / w;
      const double e = exp(-z * z);
      u += c[k] * e;
      u1 += c[k] * e * (-2.0 * z / w);
      u2 += c[k] * e * (4.0 * z * z - 2.0) / (w * w);
    }
  } else {
    const double la = p.jas[0];
    const double a = (la > 0 ? la + log1p(exp(-la)) : log1p(exp(la))) + 0.05;
    const double al = same ? p.cusp / 3.0 : p.cusp;
    const double e = exp(-r / a);
    u = al * a * (1.0 - e);
    u1 = al * e;
    u2 = -al / a * e;
  }
}

// Pair distance of the Jastrow and its derivatives with respect to r_a (r = r_a - r_b), 2D:
//   kinds 0, 1 (psiformer.py:164-170): the minimum-image distance d = |r - A round(A^-1 r)|,
//     grad d = m / d, Lap d = 2/d - |m|^2/d^3 (with the 1e-12 regularisation inside d);
//   kind 2 (no reference counterpart): the periodic chord distance of a rectangular cell,
//     d^2 = sum_i (L_i/pi)^2 sin^2(pi x_i / L_i), equal to |r| near r = 0 (the Kato cusp is kept)
//     and smooth everywhere else on the torus, so the local energy has no hidden delta terms on
//     the cell boundary (the minimum image has a slope jump there).
__device__ inline void pair_dist(const DetParams& p, double dx, double dy, double& d, double& gx, double& gy,
                                 double& lapd) {
  if (p.jastrow_kind != 2) {
    double fx = p.Ainv[0] * dx + p.Ainv[1] * dy, fy = p.Ainv[2] * dx + p.Ainv[3] * dy;
    fx -= rint(fx);
    fy -= rint(fy);
    const double mx = p.A[0] * fx + p.A[1] * fy, my = p.A[2] * fx + p.A[3] * fy;
    const double d2 = mx * mx + my * my;
    d = sqrt(d2 + 1e-12);
    gx = mx / d;
    gy = my / d;
    lapd = 2.0 / d - d2 / (d * d * d);
    return;
  }
  const double Lx = p.A[0], Ly = p.A[3];
  double sx, cx, sy, cy, s2x, c2x, s2y, c2y;
  sincos(M_PI * dx / Lx, &sx, &cx);
  sincos(M_PI * dy / Ly, &sy, &cy);
  sincos(2.0 * M_PI * dx / Lx, &s2x, &c2x);
  sincos(2.0 * M_PI * dy / Ly, &s2y, &c2y);
  const double D = (Lx / M_PI) * (Lx / M_PI) * sx * sx + (Ly / M_PI) * (Ly / M_PI) * sy * sy;
  const double gDx = (Lx / M_PI) * s2x, gDy = (Ly / M_PI) * s2y;
  const double lapD = 2.0 * c2x + 2.0 * c2y;
  d = sqrt(D + 1e-12);
  gx = gDx / (2.0 * d);
  gy = gDy / (2.0 * d);
  lapd = lapD / (2.0 * d) - (gDx * gDx + gDy * gDy) / (4.0 * d * d * d);
}

template <typename T>
__global__ __launch_bounds__(256) void k_det_jet(JetGeom j, DetParams p, const double* __restrict__ R,
                                                 const int32_t* __restrict__ S, int64_t w0,
                                                 const T* __restrict__ bf, double* __restrict__ out_lp,
                                                 double* __restrict__ out_tan, double* __restrict__ out_lap) {
  extern __shared__ double shd[];
  const int N = j.N, NN = N * N, K = p.K, nt = j.nt;
  const int tid = threadIdx.x, nthr = blockDim.x;
  cd* env = reinterpret_cast<cd*>(shd);
  cd* dex = env + NN;
  cd* dey = dex + NN;
  cd* lenv = dey + NN;
  cd* M = lenv + NN;
  cd* Mi = M + NN;
  cd* Ac = Mi + NN;
  cd* res = Ac + NN;                      // per det: [logdet, tan(nt), lap]
  double* pos = reinterpret_cast<double*>(res + (size_t)K * (nt + 2));
  int* spin = reinterpret_cast<int*>(pos + 2 * N);
  int* piv = spin + N;
  int* flag = piv + N;                    // [0] swap parity, [1] singular
  double* red = reinterpret_cast<double*>(flag + 4);

  const int64_t cfg = blockIdx.x;
  int64_t w;
  int flip = -1;
  if (j.mode == 2) {
    w = w0 + cfg / N;
    flip = (int)(cfg % N);
  } else {
    w = w0 + cfg;
  }
  for (int i = tid; i < N; i += nthr) {
    pos[2 * i] = R[(w * N + i) * 2];
    pos[2 * i + 1] = R[(w * N + i) * 2 + 1];
    int s = S[w * N + i];
    spin[i] = (i == flip) ? 1 - s : s;
  }
  __syncthreads();
  // ---- envelope and its derivatives (column i depends on r_i only)
  for (int e = tid; e < NN; e += nthr) {
    const int a = e / N, i = e % N;
    cd v = mk(0, 0), gx = mk(0, 0), gy = mk(0, 0), lp = mk(0, 0);
    for (int g = 0; g < p.n_g; ++g) {
      const double Gx = p.gs[2 * g], Gy = p.gs[2 * g + 1];
      double sn, cs;
      sincos(Gx * pos[2 * i] + Gy * pos[2 * i + 1], &sn, &cs);
      const double* wp = p.wenv + ((size_t)(a * p.n_g + g) * 2 + spin[i]) * 2;
      const cd t = mk(wp[0], wp[1]) * mk(cs, sn);
      v = v + t;
      gx = gx + t * mk(0, Gx);
      gy = gy + t * mk(0, Gy);
      lp = lp + (-(Gx * Gx + Gy * Gy)) * t;
    }
    env[e] = v;
    dex[e] = gx;
    dey[e] = gy;
    lenv[e] = lp;
  }
  __syncthreads();
  const int C = j.C;
  const int64_t KN = (int64_t)K * N;
  const int64_t rowlen = 2 * KN;
  const int64_t chs = (int64_t)N * rowlen;
  const T* bfc0 = bf + cfg * C * chs;
  auto bfval = [&](int c, int k, int a, int i) -> cd {
    const T* r = bfc0 + c * chs + (int64_t)i * rowlen;
    return mk((double)r[k * N + a], (double)r[KN + k * N + a]);
  };
  auto tan_elec = [&](int c, int& e, int& b) {
    if (j.mode == 2) {
      e = flip;
      b = c - 1;
    } else {
      e = (c - 1) >> 1;
      b = (c - 1) & 1;
    }
  };

  for (int k = 0; k < K; ++k) {
    for (int e = tid; e < NN; e += nthr) {
      const int a = e / N, i = e % N;
      M[e] = env[e] * (mk(1, 0) + bfval(0, k, a, i));
    }
    __syncthreads();
    wg_lu(M, N, piv, flag);
    if (tid == 0) res[(size_t)k * (nt + 2)] = lu_logdet(M, N, flag[0]);
    if (nt == 0 && !j.has_lap) {
      __syncthreads();
      continue;
    }
    wg_inverse(M, Mi, N, piv);
    cd hsum = mk(0, 0);
    for (int c = 1; c <= nt; ++c) {
      int ec, bc;
      tan_elec(c, ec, bc);
      // dM_c in Ac
      for (int e = tid; e < NN; e += nthr) {
        const int a = e / N, i = e % N;
        cd v = env[e] * bfval(c, k, a, i);
        if (i == ec) v = v + (bc == 0 ? dex[e] : dey[e]) * (mk(1, 0) + bfval(0, k, a, i));
        Ac[e] = v;
      }
      __syncthreads();
      if (!j.has_lap) {
        double tr = 0, ti = 0;
        for (int e = tid; e < NN; e += nthr) {
          const int i = e / N, a = e % N;
          const cd t = Mi[i * N + a] * Ac[a * N + i];
          tr += t.re;
          ti += t.im;
        }
        block_sum2(tr, ti, red);
        if (tid == 0) res[(size_t)k * (nt + 2) + c] = mk(tr, ti);
      } else {
        // A = Mi dM_c (N x N); each thread its elements, kept in registers for the trace terms
        cd av[4];
        int cnt = 0;
        for (int e = tid; e < NN; e += nthr, ++cnt) {
          const int i = e / N, jj = e % N;
          cd s = mk(0, 0);
          for (int a = 0; a < N; ++a) s = s + Mi[i * N + a] * Ac[a * N + jj];
          av[cnt] = s;
        }
        __syncthreads();
        cnt = 0;
        for (int e = tid; e < NN; e += nthr, ++cnt) Ac[e] = av[cnt];
        __syncthreads();
        double tr = 0, ti = 0, hr = 0, hi = 0;
        cnt = 0;
        for (int e = tid; e < NN; e += nthr, ++cnt) {
          const int i = e / N, jj = e % N;
          if (i == jj) {
            tr += av[cnt].re;
            ti += av[cnt].im;
          }
          const cd t = av[cnt] * Ac[jj * N + i];
          hr += t.re;
          hi += t.im;
        }
        block_sum2(tr, ti, red);
        block_sum2(hr, hi, red);
        if (tid == 0) res[(size_t)k * (nt + 2) + c] = mk(tr, ti);
        hsum = hsum + mk(hr, hi);
      }
    }
    if (j.has_lap) {
      const int cl = nt + 1;
      for (int e = tid; e < NN; e += nthr) {
        const int a = e / N, i = e % N;
        const cd one_bf = mk(1, 0) + bfval(0, k, a, i);
        cd v = lenv[e] * one_bf + env[e] * bfval(cl, k, a, i);
        v = v + 2.0 * (dex[e] * bfval(1 + 2 * i, k, a, i) + dey[e] * bfval(2 + 2 * i, k, a, i));
        Ac[e] = v;
      }
      __syncthreads();
      double tr = 0, ti = 0;
      for (int e = tid; e < NN; e += nthr) {
        const int i = e / N, a = e % N;
        const cd t = Mi[i * N + a] * Ac[a * N + i];
        tr += t.re;
        ti += t.im;
      }
      block_sum2(tr, ti, red);
      if (tid == 0) res[(size_t)k * (nt + 2) + nt + 1] = mk(tr, ti) - hsum;
    }
    __syncthreads();
  }
  // ---- log-sum-exp over determinants + Jastrow, by thread 0 (K x nt work)
  if (tid == 0) {
    cd L;
    double amax = -INFINITY;
    for (int k = 0; k < K; ++k) {
      cd z = res[(size_t)k * (nt + 2)];
      if (K > 1) z = z + mk(p.det_logw[2 * k], p.det_logw[2 * k + 1]);
      res[(size_t)k * (nt + 2)] = z;
      amax = z.re > amax ? z.re : amax;
    }
    if (K == 1) {
      L = res[0];
    } else {
      if (!isfinite(amax)) amax = 0.0;
      cd s = mk(0, 0);
      for (int k = 0; k < K; ++k) s = s + cexp_(res[(size_t)k * (nt + 2)] - mk(amax, 0));
      L = clog_(s) + mk(amax, 0);
      // weights p_k = exp(z_k - L)
      for (int k = 0; k < K; ++k) res[(size_t)k * (nt + 2)] = cexp_(res[(size_t)k * (nt + 2)] - L);
    }
    cd lap = mk(0, 0), g2 = mk(0, 0);
    for (int c = 1; c <= nt; ++c) {
      cd g = mk(0, 0);
      for (int k = 0; k < K; ++k) {
        const cd pk = (K == 1) ? mk(1, 0) : res[(size_t)k * (nt + 2)];
        const cd gk = res[(size_t)k * (nt + 2) + c];
        g = g + pk * gk;
        if (j.has_lap) lap = lap + pk * (gk * gk);
      }
      red[2 + 2 * c] = g.re;  // stash (red has room: sized for 2 + 2 (nt+1))
      red[3 + 2 * c] = g.im;
      g2 = g2 + g * g;
    }
    if (j.has_lap) {
      for (int k = 0; k < K; ++k) {
        const cd pk = (K == 1) ? mk(1, 0) : res[(size_t)k * (nt + 2)];
        lap = lap + pk * res[(size_t)k * (nt + 2) + nt + 1];
      }
      lap = lap - g2;
    }
    // Jastrow
    if (p.use_jastrow) {
      double J = 0, lJ = 0;
      for (int c = 1; c <= nt; ++c) red[2 + 2 * (nt + 1) + c] = 0;  // dJ per tangent
      for (int a = 0; a < N; ++a)
        for (int b = a + 1; b < N; ++b) {
          const double dx = pos[2 * a] - pos[2 * b], dy = pos[2 * a + 1] - pos[2 * b + 1];
          double r, grx, gry, lapr;
          pair_dist(p, dx, dy, r, grx, gry, lapr);
          double u, u1, u2;
          jastrow_u(p, spin[a] == spin[b], r, u, u1, u2);
          J += u;
          lJ += 2.0 * (u2 * (grx * grx + gry * gry) + u1 * lapr);
          for (int c = 1; c <= nt; ++c) {
            int ec, bc;
            tan_elec(c, ec, bc);
            const double comp = (bc == 0 ? grx : gry) * u1;
            if (ec == a) red[2 + 2 * (nt + 1) + c] += comp;
            if (ec == b) red[2 + 2 * (nt + 1) + c] -= comp;
          }
        }
      L = L + mk(J, 0);
      if (j.has_lap) lap = lap + mk(lJ, 0);
      for (int c = 1; c <= nt; ++c) red[2 + 2 * c] += red[2 + 2 * (nt + 1) + c];
    }
    out_lp[2 * cfg] = L.re;
    out_lp[2 * cfg + 1] = L.im;
    for (int c = 1; c <= nt; ++c) {
      out_tan[(cfg * nt + (c - 1)) * 2] = red[2 + 2 * c];
      out_tan[(cfg * nt + (c - 1)) * 2 + 1] = red[3 + 2 * c];
    }
    if (j.has_lap && out_lap) {
      out_lap[2 * cfg] = lap.re;
      out_lap[2 * cfg + 1] = lap.im;
    }
  }
}

// ------------------------------------------------------------------ wave-level head for N <= 16
// plane waves whose phases the N <= 16 head tabulates in LDS (more are evaluated per entry)
constexpr int kPhaseMax = 64;
// One workgroup of 4 waves per configuration; the envelope jet is built once in LDS, then each
// wave takes determinants k = wave, wave + 4, ... entirely in registers.  Matrices are padded to
// 16 x 16 with an identity block (same determinant, inverse and traces).
//
// Two register layouts of a 16 x 16 complex matrix X, lane = 16 g + l (l = lane % 16):
//   row chunk  ("A"): slot s holds X[l][4 g + s]
//   column chunk ("B"): slot s holds X[4 g + s][l]
// M and its inverse live in the row-chunk layout, which is also the A operand of
// v_mfma_f64_16x16x4f64 with k index 4 g + s in step s; the tangent matrices dM live in the
// column-chunk layout, the matching B operand.  Hence tr(Mi dM) is a lane-local dot product
// summed over the wave, and tr((Mi dM)^2) = sum_ij A_ij A_ji pairs each accumulator slot of
// A = Mi dM with the same slot of (dM^T Mi^T) = A^T, obtained from the same registers with the
// MFMA operands swapped.
//
// LU with partial pivoting (largest |re| + |im|, first index on ties, as the block kernel
// above) moves rows with cross-lane shuffles; `orig` follows the original index of the row at
// each position, so P I is known without replaying the swaps.
// x, opaque to the optimiser: per-slot addresses built from it inside a loop stay inside the
// loop instead of being hoisted (and held in registers) across the determinant's whole life
__device__ __forceinline__ int opaque_i(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

__device__ __forceinline__ cd shfl_cd(cd v, int src) { return mk(__shfl(v.re, src), __shfl(v.im, src)); }

// DPP moves inside each 16-lane row (VALU, no LDS traffic): row_newbcast:L (0x150 + L) gives every
// lane the value of lane L of its row, row_ror:K (0x120 + K) the value of lane (i - K) mod 16.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = dpp_i<CTRL>((int)b), hi = dpp_i<CTRL>((int)(b >> 32));
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
template <int L>
__device__ __forceinline__ cd bcast16(cd v) {
  return mk(dpp_d<0x150 + L>(v.re), dpp_d<0x150 + L>(v.im));
}
// the value of one lane, wave-uniform
__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
// Sum over the wave: pairwise within quads (quad_perm [1,0,3,2], [2,3,0,1]), then quad with
// quad (row_half_mirror: lane i with 7 - i), half-row with half-row (row_mirror: i with 15 - i),
// all as DPP, then across the four rows with two shuffles.  Every step adds two lanes' values
// in both lanes, so all 64 lanes end with the same bits.
__device__ __forceinline__ double wave_sum_d(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}
__device__ __forceinline__ cd wave_sum_cd(cd v) { return mk(wave_sum_d(v.re), wave_sum_d(v.im)); }
// product over the wave, same pattern
__device__ __forceinline__ cd wave_prod_cd(cd v) {
  v = v * mk(dpp_d<0xB1>(v.re), dpp_d<0xB1>(v.im));
  v = v * mk(dpp_d<0x4E>(v.re), dpp_d<0x4E>(v.im));
  v = v * mk(dpp_d<0x141>(v.re), dpp_d<0x141>(v.im));
  v = v * mk(dpp_d<0x140>(v.re), dpp_d<0x140>(v.im));
  v = v * mk(__shfl_xor(v.re, 16), __shfl_xor(v.im, 16));
  v = v * mk(__shfl_xor(v.re, 32), __shfl_xor(v.im, 32));
  return v;
}

struct Lu16 {
  cd m[4];      // row-chunk layout; after factor(): L below the diagonal (unit), U on and above
  int orig;     // original row index of the row now at this lane's position
  int parity;   // parity of the row swaps
  bool zero_pivot;
};

// slot s (0..3, wave-uniform at run time) of the 4 slots: an extract from a vector value, so the
// slots stay in registers (an array indexed at run time would be moved to scratch)
typedef double d8v __attribute__((ext_vector_type(8)));
__device__ __forceinline__ cd pick4(const cd (&m)[4], int s) {
  d8v v;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v[2 * t] = m[t].re;
    v[2 * t + 1] = m[t].im;
  }
  return mk(v[2 * s], v[2 * s + 1]);
}

// Pivot step Q (compile time, so that row broadcasts are DPP row_newbcast:Q): the column and the
// row swap move across rows with ds_bpermute; the pivot search (rotations), the pivot and the pivot
// row stay inside each 16-lane row.  Branch-free: the matrix keeps one set of registers.
template <int Q>
__device__ __forceinline__ void lu16_step(Lu16& f, int lane) {
  constexpr int SQ = Q & 3, GQ = Q >> 2;
  const int r = lane & 15, g = lane >> 4;
  cd cq = shfl_cd(f.m[SQ], r + 16 * GQ);  // X[r][Q]
  double v = r >= Q ? fabs(cq.re) + fabs(cq.im) : -1.0;
  int br = r;
#define NQS_LU_MAX(K)                                                        \
  {                                                                          \
    const double ov = dpp_d<0x120 + K>(v);                                   \
    const int obr = dpp_i<0x120 + K>(br);                                    \
    const bool take = ov > v || (ov == v && obr < br);                       \
    v = take ? ov : v;                                                       \
    br = take ? obr : br;                                                    \
  }
  NQS_LU_MAX(1)
  NQS_LU_MAX(2)
  NQS_LU_MAX(4)
  NQS_LU_MAX(8)
#undef NQS_LU_MAX
  const int p = __builtin_amdgcn_readfirstlane(br);
  f.zero_pivot = f.zero_pivot || __builtin_amdgcn_readfirstlane(v == 0.0 ? 1 : 0) != 0;
  f.parity ^= p != Q ? 1 : 0;
  const int src = (r == Q ? p : (r == p ? Q : r)) + 16 * g;  // p == Q: every lane onto itself
#pragma unroll
  for (int s = 0; s < 4; ++s) f.m[s] = shfl_cd(f.m[s], src);
  f.orig = __shfl(f.orig, src);
  cq = shfl_cd(cq, src);
  const cd pv = bcast16<Q>(cq);
  const bool zero = pv.re == 0.0 && pv.im == 0.0;
  const cd l = (r > Q && !zero) ? cdiv_(cq, pv) : mk(0, 0);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const cd rq = bcast16<Q>(f.m[s]);
    const int col = 4 * g + s;
    if (r > Q) {
      if (col == Q) f.m[s] = l;
      else if (col > Q) f.m[s] = f.m[s] - l * rq;
    }
  }
}

template <int Q>
struct Lu16Steps {
  static __device__ __forceinline__ void run(Lu16& f, int lane) {
    lu16_step<Q>(f, lane);
    Lu16Steps<Q + 1>::run(f, lane);
  }
};
template <>
struct Lu16Steps<16> {
  static __device__ __forceinline__ void run(Lu16&, int) {}
};

__device__ __forceinline__ void lu16_factor(Lu16& f, int lane) {
  f.orig = lane & 15;
  f.parity = 0;
  f.zero_pivot = false;
  Lu16Steps<0>::run(f, lane);
}

// log det from the factors (principal branch), as lu_logdet for the first N pivots.
__device__ __forceinline__ cd lu16_logdet(const Lu16& f, int lane, int N) {
  const int r = lane & 15, g = lane >> 4;
  const bool mine = (r >> 2) == g && r < N;
  const cd u = pick4(f.m, r & 3);
  const double au = mine ? cabs_(u) : 1.0;
  double la = wave_sum_d(log(au));
  cd ph = wave_prod_cd(mine && au > 0 ? mk(u.re / au, u.im / au) : mk(1, 0));
  if (f.parity) ph = mk(-ph.re, -ph.im);
  return mk(la + log(hypot(ph.re, ph.im)), atan2(ph.im, ph.re));
}

// Mi = U^-1 L^-1 P in the row-chunk layout.  Step Q at compile time: row Q of y is a DPP
// broadcast, the column of L or U one ds_bpermute, U[Q][Q] one readlane.
template <int Q>
struct Lu16Inv {
  static __device__ __forceinline__ void fwd(const Lu16& f, int lane, cd (&y)[4]) {
    const int r = lane & 15;
    const cd lq = shfl_cd(f.m[Q & 3], r + 16 * (Q >> 2));  // L[r][Q]
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const cd yq = bcast16<Q>(y[s]);
      if (r > Q) y[s] = y[s] - lq * yq;
    }
    Lu16Inv<Q + 1>::fwd(f, lane, y);
  }
  static __device__ __forceinline__ void bwd(const Lu16& f, int lane, cd (&y)[4]) {
    const int r = lane & 15;
    const cd mq = f.m[Q & 3];
    constexpr int UL = Q + 16 * (Q >> 2);  // lane of U[Q][Q]
    const cd uqq = mk(readlane_d(mq.re, UL), readlane_d(mq.im, UL));
    if (r == Q)
#pragma unroll
      for (int s = 0; s < 4; ++s) y[s] = cdiv_(y[s], uqq);
    const cd urq = shfl_cd(mq, r + 16 * (Q >> 2));  // U[r][Q]
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const cd zq = bcast16<Q>(y[s]);
      if (r < Q) y[s] = y[s] - urq * zq;
    }
    Lu16Inv<Q - 1>::bwd(f, lane, y);
  }
};
template <>
struct Lu16Inv<16> {
  static __device__ __forceinline__ void fwd(const Lu16&, int, cd (&)[4]) {}
};
template <>
struct Lu16Inv<-1> {
  static __device__ __forceinline__ void bwd(const Lu16&, int, cd (&)[4]) {}
};

__device__ __forceinline__ void lu16_inverse(const Lu16& f, int lane, cd (&y)[4]) {
  const int g = lane >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) y[s] = mk(4 * g + s == f.orig ? 1.0 : 0.0, 0);
  Lu16Inv<0>::fwd(f, lane, y);
  Lu16Inv<15>::bwd(f, lane, y);
}

template <typename T>
__device__ __forceinline__ cd bf_at(const T* __restrict__ bfc, int64_t rowlen, int64_t KN, int N, int k, int a, int i) {
  const bool ok = a < N && i < N;
  const int ac = a < N ? a : N - 1, ic = i < N ? i : N - 1;
  const T* rp = bfc + (int64_t)ic * rowlen + (int64_t)k * N + ac;
  const double re = (double)rp[0], im = (double)rp[KN];
  return ok ? mk(re, im) : mk(0, 0);
}

// Log-sum-exp over the determinants, tangents and Laplacian of log psi, Jastrow: one wave,
// from res[k * rs_ + c] (c = 0 log det, 1..nt tangents, nt + 1 Laplacian term).
template <int KIND>
__device__ __forceinline__ void det_head_tail(const JetGeom& j, const DetParams& p, const cd* res, int rs_,
                                              const double* pos, const int* spin, int64_t cfg, int flip,
                                              double* __restrict__ out_lp, double* __restrict__ out_tan,
                                              double* __restrict__ out_lap) {
  const int N = j.N, K = p.K, nt = j.nt;
  const int lane = threadIdx.x & 63;
  // ---- log-sum-exp over determinants, tangents and Laplacian of log psi, Jastrow
  cd L;
  double amax = -INFINITY;
  for (int k = 0; k < K; ++k) {
    cd z = res[k * rs_];
    if (K > 1) z = z + mk(p.det_logw[2 * k], p.det_logw[2 * k + 1]);
    amax = z.re > amax ? z.re : amax;
  }
  if (K == 1) {
    L = res[0];
  } else {
    if (!isfinite(amax)) amax = 0.0;
    cd s = mk(0, 0);
    for (int k = 0; k < K; ++k) s = s + cexp_(res[k * rs_] + mk(p.det_logw[2 * k], p.det_logw[2 * k + 1]) - mk(amax, 0));
    L = clog_(s) + mk(amax, 0);
  }
  auto pk_of = [&](int k) -> cd {
    return K == 1 ? mk(1, 0) : cexp_(res[k * rs_] + mk(p.det_logw[2 * k], p.det_logw[2 * k + 1]) - L);
  };
  // tangent c = lane + 1 (nt <= 2N <= 32)
  const int c = lane + 1;
  const bool has_c = c <= nt;
  cd gsum = mk(0, 0), lsum = mk(0, 0);
  if (has_c || KIND == 2)
    for (int k = 0; k < K; ++k) {
      const cd pk = pk_of(k);
      if (has_c) {
        const cd gk = res[k * rs_ + c];
        gsum = gsum + pk * gk;
        if constexpr (KIND == 2) lsum = lsum + pk * (gk * gk);
      }
    }
  cd lap = mk(0, 0);
  if constexpr (KIND == 2) {
    cd lk = mk(0, 0);
    for (int k = 0; k < K; ++k) lk = lk + pk_of(k) * res[k * rs_ + nt + 1];
    const cd g2 = wave_sum_cd(has_c ? gsum * gsum : mk(0, 0));
    lap = wave_sum_cd(lsum) + lk - g2;
  }
  if (p.use_jastrow) {
    // lane a (< N) sums its pairs: gradient of J wrt r_a over all partners, J and its
    // Laplacian over partners b > a
    const int a = lane;
    double J = 0, lJ = 0, gxa = 0, gya = 0;
    if (a < N) {
      for (int b = 0; b < N; ++b) {
        if (b == a) continue;
        const double dx = pos[2 * a] - pos[2 * b], dy = pos[2 * a + 1] - pos[2 * b + 1];
        double r, grx, gry, lapr;
        pair_dist(p, dx, dy, r, grx, gry, lapr);
        double u, u1, u2;
        jastrow_u(p, spin[a] == spin[b], r, u, u1, u2);
        gxa += grx * u1;
        gya += gry * u1;
        if (b > a) {
          J += u;
          lJ += 2.0 * (u2 * (grx * grx + gry * gry) + u1 * lapr);
        }
      }
    }
    J = wave_sum(J);
    lJ = wave_sum(lJ);
    int ec = 0, dc = 0;
    if (has_c) {
      if (j.mode == 2) {
        ec = flip;
        dc = c - 1;
      } else {
        ec = (c - 1) >> 1;
        dc = (c - 1) & 1;
      }
    }
    const double gxe = __shfl(gxa, ec), gye = __shfl(gya, ec);
    gsum = gsum + mk(dc == 0 ? gxe : gye, 0);
    L = L + mk(J, 0);
    lap = lap + mk(lJ, 0);
  }
  if (lane == 0) {
    out_lp[2 * cfg] = L.re;
    out_lp[2 * cfg + 1] = L.im;
    if (KIND == 2 && out_lap) {
      out_lap[2 * cfg] = lap.re;
      out_lap[2 * cfg + 1] = lap.im;
    }
  }
  if (has_c) {
    out_tan[(cfg * nt + (c - 1)) * 2] = gsum.re;
    out_tan[(cfg * nt + (c - 1)) * 2 + 1] = gsum.im;
  }
}

// KIND (compile time, so that each variant keeps only the registers it needs): 0 log psi only,
// 1 with tangents, 2 with tangents and the Laplacian.  (Two determinants factored at once per wave,
// their pivot steps interleaved, measured 0.65 ms against 0.48 ms for the value-only head at 8192
// configurations: the doubled registers halve the waves in flight.)
template <typename T, int KIND>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(KIND == 0 ? 4 : (KIND == 1 ? 3 : 2))))
void k_det_jet16(JetGeom j, DetParams p, const double* __restrict__ R,
                                                   const int32_t* __restrict__ S, int64_t w0,
                                                   const T* __restrict__ bf, double* __restrict__ out_lp,
                                                   double* __restrict__ out_tan, double* __restrict__ out_lap) {
  // the envelope's derivative tables exist only where the variant reads them
  __shared__ cd env[256], dex[KIND > 0 ? 256 : 1], dey[KIND > 0 ? 256 : 1], lenv[KIND == 2 ? 256 : 1];
  __shared__ cd ph[kPhaseMax * 16];  // exp(i G_g . r_i), [g][i], when n_g <= kPhaseMax
  __shared__ double pos[32];
  __shared__ int spin[16];
  extern __shared__ cd res[];  // [K][nt + 2]: logdet, tangents, Laplacian
  const int N = j.N, K = p.K, nt = j.nt;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  const int64_t cfg = blockIdx.x;
  int64_t w;
  int flip = -1;
  if (j.mode == 2) {
    w = w0 + cfg / N;
    flip = (int)(cfg % N);
  } else {
    w = w0 + cfg;
  }
  if (tid < N) {
    pos[2 * tid] = R[(w * N + tid) * 2];
    pos[2 * tid + 1] = R[(w * N + tid) * 2 + 1];
    const int s = S[w * N + tid];
    spin[tid] = (tid == flip) ? 1 - s : s;
  }
  __syncthreads();
  // the phases depend on (g, i) only: evaluated once here instead of once per orbital
  const bool tab = p.n_g <= kPhaseMax;
  if (tab) {
    for (int e = tid; e < p.n_g * 16; e += blockDim.x) {
      const int gg = e >> 4, i = e & 15;
      double sn = 0, cs = 0;
      if (i < N) sincos(p.gs[2 * gg] * pos[2 * i] + p.gs[2 * gg + 1] * pos[2 * i + 1], &sn, &cs);
      ph[e] = mk(cs, sn);
    }
    __syncthreads();
  }
  for (int e = tid; e < 256; e += blockDim.x) {  // envelope jet, entry (a, i) = (e / 16, e % 16)
    const int a = e >> 4, i = e & 15;
    cd v = mk(0, 0), gx = mk(0, 0), gy = mk(0, 0), lp = mk(0, 0);
    if (a < N && i < N) {
      for (int gg = 0; gg < p.n_g; ++gg) {
        const double Gx = p.gs[2 * gg], Gy = p.gs[2 * gg + 1];
        double sn, cs;
        if (tab) {
          cs = ph[gg * 16 + i].re;
          sn = ph[gg * 16 + i].im;
        } else {
          sincos(Gx * pos[2 * i] + Gy * pos[2 * i + 1], &sn, &cs);
        }
        const double* wp = p.wenv + ((size_t)(a * p.n_g + gg) * 2 + spin[i]) * 2;
        const cd t = mk(wp[0], wp[1]) * mk(cs, sn);
        v = v + t;
        if constexpr (KIND > 0) {
          gx = gx + t * mk(0, Gx);
          gy = gy + t * mk(0, Gy);
        }
        if constexpr (KIND == 2) lp = lp + (-(Gx * Gx + Gy * Gy)) * t;
      }
    }
    env[e] = v;
    if constexpr (KIND > 0) {
      dex[e] = gx;
      dey[e] = gy;
    }
    if constexpr (KIND == 2) lenv[e] = lp;
  }
  __syncthreads();
  const int C = j.C;
  const int64_t KN = (int64_t)K * N, rowlen = 2 * KN, chs = (int64_t)N * rowlen;
  const T* bfc0 = bf + cfg * C * chs;
  const int l16 = lane & 15, g = lane >> 4;
  const int rs_ = nt + 2;
  for (int k = wave; k < K; k += nwaves) {
    Lu16 f;
    const int l16 = opaque_i(lane & 15), g = opaque_i(lane >> 4);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int a = l16, i = 4 * g + s;
      const cd b = bf_at(bfc0, rowlen, KN, N, k, a, i);
      const cd e = env[a * 16 + i];
      f.m[s] = (a < N && i < N) ? e * (mk(1, 0) + b) : mk(a == i ? 1.0 : 0.0, 0);
    }
    lu16_factor(f, lane);
    const cd ld = lu16_logdet(f, lane, N);
    if (lane == 0) res[k * rs_] = ld;
    if constexpr (KIND == 0) continue;
    cd mi[4];
    lu16_inverse(f, lane, mi);
    // (1 + bf) in the column-chunk layout
    cd ob[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) ob[s] = mk(1, 0) + bf_at(bfc0, rowlen, KN, N, k, 4 * g + s, l16);
    cd hsum = mk(0, 0);
    // Lap M's cross term 2 sum_c denv_c dbf_c lives in column i of tangent c = 1 + 2 i + d, which the
    // tangent loop below already reads: summed there
    cd lapx = mk(0, 0);
    for (int c = 1; c <= nt; ++c) {
      int ec, dc;
      if (j.mode == 2) {
        ec = flip;
        dc = c - 1;
      } else {
        ec = (c - 1) >> 1;
        dc = (c - 1) & 1;
      }
      const T* bfc = bfc0 + c * chs;
      cd dm[4];
      cd tr = mk(0, 0);
      const int l16c = opaque_i(lane & 15), gc = opaque_i(lane >> 4);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int a = 4 * gc + s, i = l16c, e_ = a * 16 + i;
        const cd bv = bf_at(bfc, rowlen, KN, N, k, a, i);
        cd v = env[e_] * bv;
        if (i == ec && a < N) {
          const cd de = dc == 0 ? dex[e_] : dey[e_];
          v = v + de * ob[s];
          if constexpr (KIND == 2) lapx = lapx + mi[s] * (2.0 * (de * bv));
        }
        dm[s] = v;
        tr = tr + mi[s] * v;
      }
      tr = wave_sum_cd(tr);
      if (lane == 0) res[k * rs_ + c] = tr;
      if constexpr (KIND == 2) {
        d4v are = (d4v)0.0, aim = (d4v)0.0, tre = (d4v)0.0, tim = (d4v)0.0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          are = __builtin_amdgcn_mfma_f64_16x16x4f64(mi[s].re, dm[s].re, are, 0, 0, 0);
          are = __builtin_amdgcn_mfma_f64_16x16x4f64(-mi[s].im, dm[s].im, are, 0, 0, 0);
          aim = __builtin_amdgcn_mfma_f64_16x16x4f64(mi[s].re, dm[s].im, aim, 0, 0, 0);
          aim = __builtin_amdgcn_mfma_f64_16x16x4f64(mi[s].im, dm[s].re, aim, 0, 0, 0);
          tre = __builtin_amdgcn_mfma_f64_16x16x4f64(dm[s].re, mi[s].re, tre, 0, 0, 0);
          tre = __builtin_amdgcn_mfma_f64_16x16x4f64(-dm[s].im, mi[s].im, tre, 0, 0, 0);
          tim = __builtin_amdgcn_mfma_f64_16x16x4f64(dm[s].re, mi[s].im, tim, 0, 0, 0);
          tim = __builtin_amdgcn_mfma_f64_16x16x4f64(dm[s].im, mi[s].re, tim, 0, 0, 0);
        }
        cd h = mk(0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) h = h + mk(are[r], aim[r]) * mk(tre[r], tim[r]);
        hsum = hsum + wave_sum_cd(h);
      }
    }
    if constexpr (KIND == 2) {
      const int cl = nt + 1;
      const T* bfl = bfc0 + cl * chs;
      cd tr = mk(0, 0);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int a = 4 * g + s, i = l16, e_ = a * 16 + i;
        const cd v = lenv[e_] * ob[s] + env[e_] * bf_at(bfl, rowlen, KN, N, k, a, i);
        tr = tr + mi[s] * v;
      }
      tr = wave_sum_cd(tr + lapx);
      if (lane == 0) res[k * rs_ + cl] = tr - hsum;
    }
  }
  __syncthreads();
  if (wave != 0) return;
  det_head_tail<KIND>(j, p, res, rs_, pos, spin, cfg, flip, out_lp, out_tan, out_lap);
}

// ------------------------------------------------------------------ row-per-lane heads, N <= 32
// The value-only pass (KIND 0) and the spin-flip pass (KIND 1: log psi and the two tangents of the
// flipped electron) with one determinant per group of NR lanes: NR = 16 for N <= 16 (four
// determinants per wave), NR = 32 for 16 < N <= 32 (two per wave); lane r of the group holds row r
// of M (orbital r; electron i in slot i), NR complex entries in registers.  Rows never move.  The
// pivot of step q is the lane with the largest |re| + |im| in column q among the rows not yet used
// (first lane on ties), found by DPP rotations inside each 16-lane row (and one exchange between
// the two rows of a 32-lane group); its row reaches the other lanes through LDS.  The pivot lanes
// p_q are packed four bits each (NR = 16) or five (NR = 32), and the sign of the permutation
// q -> p_q is the parity of its inversions.
//   KIND 0  elimination below the pivots; log det = sum of log pivot + the permutation's sign.
//   KIND 1  Gauss-Jordan in place: after step q, slot q of lane r holds E[r][p_q], E the product
//           of the row operations (E M = P with P[p_q][q] = 1), so M^-1[q][c] = E[p_q][c]: lane
//           r = p_q holds row q of M^-1 with its columns in pivot order, and
//           tr(M^-1 dM) = sum_r sum_s slot_s(r) dM[p_s][q(r)], a per-lane dot product with dM
//           gathered along the pivot order.
// Steps q >= N are skipped: the identity padding (to NR) has pivots exactly 1 and its rows and
// columns take no part in the real block.  The multiplier of a step is the row entry times the pivot's
// reciprocal conj(u) / |u|^2, as LAPACK's getf2 scales a column by 1 / u.  (The pivot search
// compares magnitudes rounded to float: entries below 1e-38 in magnitude would compare as zero.)  A zero pivot leaves
// the rows as they are (log det = -inf, as the LU of the N <= 16 head above), and the tangents of
// a singular matrix are NaN, as its substitution gives them.
template <int NR>
struct RowPk {  // p_q packed BITS bits each, PER per 32-bit word, in NW words
  static constexpr int BITS = NR == 16 ? 4 : 5, PER = NR == 16 ? 8 : 6, NW = (NR + PER - 1) / PER;
};
template <int NR>
struct RowDet {
  cd m[NR];
  cd piv;                       // this lane's pivot (1 while the lane has not been one)
  int q;                        // the step at which this lane was the pivot
  unsigned pk[RowPk<NR>::NW];   // p_q of steps PER w .. PER w + PER - 1 in pk[w]
  bool used, zero;
};

// Reductions over the NR lanes of a group, the same bits in every lane of it: DPP inside each
// 16-lane row, then (NR = 32) the two rows combined in the same operand order in both.
template <int NR = 16>
__device__ __forceinline__ double group_sum_d(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  if constexpr (NR == 32) {
    const double o = __shfl_xor(v, 16);
    v = (threadIdx.x & 16) ? o + v : v + o;
  }
  return v;
}
template <int NR = 16>
__device__ __forceinline__ cd group_sum_cd(cd v) { return mk(group_sum_d<NR>(v.re), group_sum_d<NR>(v.im)); }
template <int NR = 16>
__device__ __forceinline__ int group_sum_i(int v) {
  v += dpp_i<0xB1>(v);
  v += dpp_i<0x4E>(v);
  v += dpp_i<0x141>(v);
  v += dpp_i<0x140>(v);
  if constexpr (NR == 32) v += __shfl_xor(v, 16);
  return v;
}
template <int NR = 16>
__device__ __forceinline__ cd group_prod_cd(cd v) {
  v = v * mk(dpp_d<0xB1>(v.re), dpp_d<0xB1>(v.im));
  v = v * mk(dpp_d<0x4E>(v.re), dpp_d<0x4E>(v.im));
  v = v * mk(dpp_d<0x141>(v.re), dpp_d<0x141>(v.im));
  v = v * mk(dpp_d<0x140>(v.re), dpp_d<0x140>(v.im));
  if constexpr (NR == 32) {
    const cd o = mk(__shfl_xor(v.re, 16), __shfl_xor(v.im, 16));
    v = (threadIdx.x & 16) ? o * v : v * o;
  }
  return v;
}
template <int NR>
__device__ __forceinline__ int rowdet_pk(const RowDet<NR>& f, int s) {
  typedef RowPk<NR> K_;
  return (int)((f.pk[s / K_::PER] >> (K_::BITS * (s % K_::PER))) & (unsigned)(NR - 1));
}

// m - l x, two fused multiply-adds per component
__device__ __forceinline__ cd cfms(cd m, cd l, cd x) {
  return mk(fma(-l.re, x.re, fma(l.im, x.im, m.re)), fma(-l.re, x.im, fma(-l.im, x.re, m.im)));
}
// t + a b
__device__ __forceinline__ cd cfma(cd t, cd a, cd b) {
  return mk(fma(a.re, b.re, fma(-a.im, b.im, t.re)), fma(a.re, b.im, fma(a.im, b.re, t.im)));
}

// Step Q.  The pivot lane writes its row (slots Q.. for KIND 0, all for KIND 1) to the group's LDS
// row xrow and every lane of the group reads it back: the lanes of a wave issue LDS operations in
// program order, and the next step's write follows this step's reads in every lane.
// 1 / x for a normal x: hardware estimate and two Newton steps
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// Step Q.  Pivot search on a 32-bit key: |re| + |im| rounded to float, its log2(NR) low bits
// replaced by NR - 1 - r, so one unsigned max per DPP rotation finds the largest magnitude (to
// 2^-19 relative for NR = 16, 2^-18 for NR = 32) and, among equal keys, the first lane; rows
// already used have key 0.  The pivot lane writes its row (slots Q.. for KIND 0, all for KIND 1)
// to the group's LDS row xrow and every lane of the group reads it back: the lanes of a wave issue
// LDS operations in program order, and the next step's write follows this step's reads in every
// lane.  The row update runs in chunks of four slots so that only one chunk of the pivot row is
// held in registers at a time.
template <int Q, int KIND, int NR>
__device__ __forceinline__ void rowdet_step(RowDet<NR>& f, int r, cd* xrow) {
  constexpr unsigned LOW = NR - 1;
  const float mag = (float)(fabs(f.m[Q].re) + fabs(f.m[Q].im));
  unsigned key = f.used ? 0u : ((__float_as_uint(mag) & ~LOW) | (LOW - (unsigned)r));
  key = max(key, (unsigned)dpp_i<0x121>((int)key));
  key = max(key, (unsigned)dpp_i<0x122>((int)key));
  key = max(key, (unsigned)dpp_i<0x124>((int)key));
  key = max(key, (unsigned)dpp_i<0x128>((int)key));
  if constexpr (NR == 32) key = max(key, (unsigned)__shfl_xor((int)key, 16));
  const int p = (int)(LOW - (key & LOW));  // the same in every lane of the group
  const bool isp = r == p;
  f.pk[Q / RowPk<NR>::PER] |= (unsigned)p << (RowPk<NR>::BITS * (Q % RowPk<NR>::PER));
  constexpr int J0 = KIND == 0 ? Q : 0;
  if (isp) {
#pragma unroll
    for (int jj = J0; jj < NR; ++jj) xrow[jj] = f.m[jj];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const cd pv = xrow[Q];
  const bool zero = pv.re == 0.0 && pv.im == 0.0;
  f.zero = f.zero || zero;
  const double in = zero ? 0.0 : rcp_nr(fma(pv.re, pv.re, pv.im * pv.im));
  const cd rp = mk(pv.re * in, -pv.im * in);
  if (isp) {
    f.piv = pv;
    f.q = Q;
  }
  if constexpr (KIND == 0) {
    const cd l = (f.used || isp) ? mk(0, 0) : f.m[Q] * rp;
#pragma unroll
    for (int jj = Q + 1; jj < NR; ++jj) {
      f.m[jj] = cfms(f.m[jj], l, xrow[jj]);
      if (((jj - Q) & 3) == 0) __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    const cd fr = f.m[Q] * rp;
    const double s = isp ? 0.0 : 1.0;
    const cd c = isp ? mk(-rp.re, -rp.im) : fr;
#pragma unroll
    for (int jj = 0; jj < NR; ++jj) {
      if (jj == Q) continue;
      f.m[jj] = cfms(mk(s * f.m[jj].re, s * f.m[jj].im), c, xrow[jj]);
      if ((jj & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    f.m[Q] = isp ? rp : mk(-fr.re, -fr.im);
  }
  f.used = f.used || isp;
}

template <int Q, int KIND, int NR>
struct RowDetSteps {
  static __device__ __forceinline__ void run(RowDet<NR>& f, int r, cd* xrow, int N) {
    if (Q >= N) return;
    rowdet_step<Q, KIND, NR>(f, r, xrow);
    RowDetSteps<Q + 1, KIND, NR>::run(f, r, xrow, N);
  }
};
template <int KIND, int NR>
struct RowDetSteps<NR, KIND, NR> {
  static __device__ __forceinline__ void run(RowDet<NR>&, int, cd*, int) {}
};

// log det (principal branch of the phase) from the pivots and the parity of q -> p_q
template <int NR>
__device__ __forceinline__ cd rowdet_logdet(const RowDet<NR>& f, int r, int N) {
  int inv = 0;
  if (f.used) {
#pragma unroll
    for (int s = 0; s < NR; ++s) inv += (s < N && s > f.q && rowdet_pk(f, s) < r) ? 1 : 0;
  }
  inv = group_sum_i<NR>(inv);
  const double au = cabs_(f.piv);
  const double la = group_sum_d<NR>(log(au));
  cd ph = group_prod_cd<NR>(au > 0 ? mk(f.piv.re / au, f.piv.im / au) : mk(1, 0));
  if (inv & 1) ph = mk(-ph.re, -ph.im);
  return mk(la + log(hypot(ph.re, ph.im)), atan2(ph.im, ph.re));
}

__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// det_head_tail for KIND 0 and 1 with nt <= 2, lane k taking determinant k (K <= 64): the
// log-sum-exp weights exp(z_k - L) are evaluated once per determinant instead of once per lane.
template <int KIND>
__device__ __forceinline__ void det_tail_lanes(const JetGeom& j, const DetParams& p, const cd* res, int rs_,
                                               const double* pos, const int* spin, int64_t cfg, int flip,
                                               double* __restrict__ out_lp, double* __restrict__ out_tan) {
  const int N = j.N, K = p.K, nt = j.nt;
  const int lane = threadIdx.x & 63;
  const bool hk = lane < K;
  cd z = hk ? res[lane * rs_] : mk(0, 0);
  cd L, pk = mk(1, 0);
  if (K == 1) {
    L = res[0];
  } else {
    if (hk) z = z + mk(p.det_logw[2 * lane], p.det_logw[2 * lane + 1]);
    double amax = wave_max_d(hk ? z.re : -INFINITY);
    if (!isfinite(amax)) amax = 0.0;
    const cd s = wave_sum_cd(hk ? cexp_(z - mk(amax, 0)) : mk(0, 0));
    L = clog_(s) + mk(amax, 0);
    pk = hk ? cexp_(z - L) : mk(0, 0);
  }
  cd g1 = mk(0, 0), g2 = mk(0, 0);
  if constexpr (KIND == 1) {
    if (hk) {
      g1 = pk * res[lane * rs_ + 1];
      if (nt > 1) g2 = pk * res[lane * rs_ + 2];
    }
    if (K > 1) {
      g1 = wave_sum_cd(g1);
      g2 = wave_sum_cd(g2);
    } else {
      g1 = res[1];
      g2 = nt > 1 ? res[2] : mk(0, 0);
    }
  }
  if (p.use_jastrow) {
    // lane a (< N) sums its pairs: gradient of J wrt r_a over all partners, J over partners b > a
    const int a = lane;
    double J = 0, gxa = 0, gya = 0;
    if (a < N) {
      for (int b = 0; b < N; ++b) {
        if (b == a) continue;
        const double dx = pos[2 * a] - pos[2 * b], dy = pos[2 * a + 1] - pos[2 * b + 1];
        double r, grx, gry, lapr;
        pair_dist(p, dx, dy, r, grx, gry, lapr);
        double u, u1, u2;
        jastrow_u(p, spin[a] == spin[b], r, u, u1, u2);
        gxa += grx * u1;
        gya += gry * u1;
        if (b > a) J += u;
      }
    }
    J = wave_sum(J);
    L = L + mk(J, 0);
    if constexpr (KIND == 1) {
      // tangents 1, 2: x and y of the flipped electron (KIND 1 is the spin-flip pass)
      g1 = g1 + mk(__shfl(gxa, flip), 0);
      g2 = g2 + mk(__shfl(gya, flip), 0);
    }
  }
  if (lane == 0) {
    out_lp[2 * cfg] = L.re;
    out_lp[2 * cfg + 1] = L.im;
    if constexpr (KIND == 1) {
      out_tan[cfg * nt * 2] = g1.re;
      out_tan[cfg * nt * 2 + 1] = g1.im;
      if (nt > 1) {
        out_tan[cfg * nt * 2 + 2] = g2.re;
        out_tan[cfg * nt * 2 + 3] = g2.im;
      }
    }
  }
}

// G configurations per workgroup, determinant t = (64 / NR) wave + lane / NR of the workgroup is
// k = t % K of configuration t / K.  Dynamic LDS (rowdet_lds): env [G][NR][NR + 1] (row pitch
// NR + 1 keeps the per-lane row reads off one bank), the flipped column's d env / dx, dy
// [G][2][NR] (KIND 1), positions [G][2 NR], spins [G][NR], flip index [G], the pivot rows of the
// 16 groups [16][NR], the envelope weights [N][n_g][2] (when small) and the plane-wave phases
// [G][n_g][NR].  Results per determinant go to resg [cfg][K][nt + 2]: log det, then the tangents.
// the envelope weights W[a][g][s] are copied to LDS when they take at most this many bytes
constexpr int kRowDetWenvMax = 16384;
__host__ __device__ inline size_t rowdet_head(int NR, int G, int kind) {
  return (size_t)G * (NR * (NR + 1) + (kind == 1 ? 2 * NR : 0)) * sizeof(cd) + (size_t)G * 2 * NR * sizeof(double) +
         (size_t)G * (NR + 1) * sizeof(int);
}
__host__ __device__ inline size_t rowdet_lds(int NR, int G, int N, int n_g, int kind) {
  const size_t wb = (size_t)N * n_g * 2 * sizeof(cd);
  const size_t a = (rowdet_head(NR, G, kind) + 15) / 16 * 16 + (size_t)16 * NR * sizeof(cd) +
                   (wb <= (size_t)kRowDetWenvMax ? wb : 0);
  return a + (n_g <= kPhaseMax ? (size_t)G * n_g * NR * sizeof(cd) : 0);
}

// Shared prologue of the row-per-lane heads: positions and spins of the workgroup's ng
// configurations (the flipped electron's spin inverted in the spin-flip pass, mode 2), the
// plane-wave phases (when n_g <= kPhaseMax) and the envelope env[c][a][i] in LDS, and for KIND 1
// the flipped column's d env / dx, dy.  wtab: the envelope weights in LDS, or null to read them
// from p.wenv.  Ends with a workgroup barrier.
template <int KIND, int NR>
__device__ __forceinline__ void rowdet_envelope(const JetGeom& j, const DetParams& p, int64_t cfg0, int ng,
                                                const double* __restrict__ R, const int32_t* __restrict__ S,
                                                int64_t w0, const cd* wtab, cd* un, cd* env, cd* dcol, double* pos,
                                                int* spin, int* flp) {
  const int N = j.N, n_g = p.n_g;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const bool wlds = wtab != nullptr;
  constexpr int EP = NR + 1;  // env row pitch
  for (int e = tid; e < ng * NR; e += nthr) {
    const int c = e / NR, t = e % NR;
    const int64_t cfg = cfg0 + c;
    int64_t w;
    int flip = -1;
    if (j.mode == 2) {
      w = w0 + cfg / N;
      flip = (int)(cfg % N);
    } else {
      w = w0 + cfg;
    }
    if (t == 0) flp[c] = flip;
    if (t < N) {
      pos[c * 2 * NR + 2 * t] = R[(w * N + t) * 2];
      pos[c * 2 * NR + 2 * t + 1] = R[(w * N + t) * 2 + 1];
      const int s = S[w * N + t];
      spin[c * NR + t] = (t == flip) ? 1 - s : s;
    }
  }
  __syncthreads();
  const bool tab = n_g <= kPhaseMax;
  if (tab) {
    for (int e = tid; e < ng * n_g * NR; e += nthr) {
      const int c = e / (n_g * NR), gg = (e / NR) % n_g, i = e % NR;
      double sn = 0, cs = 0;
      if (i < N)
        sincos(p.gs[2 * gg] * pos[c * 2 * NR + 2 * i] + p.gs[2 * gg + 1] * pos[c * 2 * NR + 2 * i + 1], &sn, &cs);
      un[e] = mk(cs, sn);
    }
    __syncthreads();
  }
  for (int e = tid; e < ng * NR * NR; e += nthr) {  // envelope, entry (a, i) of configuration c
    const int c = e / (NR * NR), a = (e / NR) % NR, i = e % NR;
    const bool fc = KIND == 1 && i == flp[c];
    cd v = mk(0, 0), gx = mk(0, 0), gy = mk(0, 0);
    if (a < N && i < N) {
      const double px = pos[c * 2 * NR + 2 * i], py = pos[c * 2 * NR + 2 * i + 1];
      const int si = spin[c * NR + i];
      for (int gg = 0; gg < n_g; ++gg) {
        const double Gx = p.gs[2 * gg], Gy = p.gs[2 * gg + 1];
        double sn, cs;
        if (tab) {
          const cd z = un[(c * n_g + gg) * NR + i];
          cs = z.re;
          sn = z.im;
        } else {
          sincos(Gx * px + Gy * py, &sn, &cs);
        }
        const int we = (a * n_g + gg) * 2 + si;
        const cd t = (wlds ? wtab[we] : mk(p.wenv[2 * we], p.wenv[2 * we + 1])) * mk(cs, sn);
        v = v + t;
        if (fc) {
          gx = gx + t * mk(0, Gx);
          gy = gy + t * mk(0, Gy);
        }
      }
    }
    env[c * NR * EP + a * EP + i] = v;
    if (fc) {
      dcol[c * 2 * NR + a] = gx;
      dcol[c * 2 * NR + NR + a] = gy;
    }
  }
  __syncthreads();
}

// NR = 16: up to four waves (256 threads); NR = 32: up to eight (512 threads), two determinants each
template <typename T, int KIND, int NR>
__global__ __launch_bounds__(NR == 16 ? 256 : 512)
__attribute__((amdgpu_waves_per_eu(NR == 16 ? (KIND == 0 ? 3 : 4) : 2)))
void k_det_rows(JetGeom j, DetParams p, int G, const double* __restrict__ R, const int32_t* __restrict__ S,
                int64_t w0, const T* __restrict__ bf, cd* __restrict__ resg) {
  extern __shared__ __align__(16) char smem[];
  constexpr int EP = NR + 1, GW = 64 / NR;
  const int N = j.N, K = p.K, nt = j.nt, rs_ = nt + 2, C = j.C;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nthr = blockDim.x;
  const int64_t cfg0 = (int64_t)blockIdx.x * G;
  const int ng = (int)min((int64_t)G, j.n_cfg - cfg0);
  cd* env = reinterpret_cast<cd*>(smem);
  cd* dcol = env + G * NR * EP;
  double* pos = reinterpret_cast<double*>(dcol + (KIND == 1 ? G * 2 * NR : 0));
  int* spin = reinterpret_cast<int*>(pos + G * 2 * NR);
  int* flp = spin + G * NR;
  cd* xrows = reinterpret_cast<cd*>(smem + (rowdet_head(NR, G, KIND) + 15) / 16 * 16);
  const int n_g = p.n_g;
  const bool wlds = N * n_g * 2 * (int)sizeof(cd) <= kRowDetWenvMax;
  cd* wtab = xrows + 16 * NR;
  cd* un = wtab + (wlds ? N * n_g * 2 : 0);
  if (wlds)
    for (int e = tid; e < N * n_g * 2; e += nthr) wtab[e] = mk(p.wenv[2 * e], p.wenv[2 * e + 1]);
  const int64_t KN = (int64_t)K * N, rowlen = 2 * KN, chs = (int64_t)N * rowlen;
  const int r = lane % NR;
  const int t = GW * wave + lane / NR;
  const int cl = t / K, k = t % K;
  const bool act = cl < ng;  // the same for the NR lanes of a group
  const T* bfc0 = bf + (cfg0 + (act ? cl : 0)) * C * chs;
  // row r of bf; the value-only pass with NR = 16 loads it before the envelope is built so that
  // its latency overlaps that work (the others have no registers to spare there)
  T bre[NR], bim[NR];
  auto load_row = [&]() {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const T* q = bfc0 + (int64_t)(i < N ? i : N - 1) * rowlen + (int64_t)k * N + (r < N ? r : N - 1);
      bre[i] = act ? q[0] : (T)0;
      bim[i] = act ? q[KN] : (T)0;
    }
  };
  if constexpr (KIND == 0 && NR == 16) load_row();
  rowdet_envelope<KIND, NR>(j, p, cfg0, ng, R, S, w0, wlds ? wtab : nullptr, un, env, dcol, pos, spin, flp);
  if constexpr (KIND == 1 || NR != 16) load_row();
  if (act) {
    const cd* envc = env + cl * NR * EP;
    RowDet<NR> f;
#pragma unroll
    for (int i = 0; i < NR; ++i)
      f.m[i] = (r < N && i < N) ? envc[r * EP + i] * mk(1.0 + (double)bre[i], (double)bim[i])
                                : mk(r == i ? 1.0 : 0.0, 0);
    f.piv = mk(1, 0);
    f.q = r;
#pragma unroll
    for (int w = 0; w < RowPk<NR>::NW; ++w) f.pk[w] = 0;
    f.used = f.zero = false;
    RowDetSteps<0, KIND, NR>::run(f, r, xrows + NR * (lane / NR) + 64 * wave, N);
    const cd ld = rowdet_logdet(f, r, N);
    cd* rk = resg + ((cfg0 + cl) * K + k) * rs_;
    if (r == 0) rk[0] = ld;
    if constexpr (KIND == 1) {
      // tangents c = 1, 2: d/dx, d/dy of the flipped electron
      const int i = f.q, flip = flp[cl];
      const cd* dc = dcol + cl * 2 * NR;
      cd t1 = mk(0, 0), t2 = mk(0, 0);
      if (f.used) {
#pragma unroll
        for (int s = 0; s < NR; ++s) {
          if (s >= N) continue;
          const int a = rowdet_pk(f, s);
          const cd e = envc[a * EP + i];
          cd d1 = e * bf_at(bfc0 + chs, rowlen, KN, N, k, a, i);
          cd d2 = e * bf_at(bfc0 + 2 * chs, rowlen, KN, N, k, a, i);
          if (i == flip) {
            const cd ob = mk(1, 0) + bf_at(bfc0, rowlen, KN, N, k, a, i);
            d1 = d1 + dc[a] * ob;
            d2 = d2 + dc[NR + a] * ob;
          }
          t1 = cfma(t1, f.m[s], d1);
          t2 = cfma(t2, f.m[s], d2);
        }
      }
      t1 = group_sum_cd<NR>(t1);
      t2 = group_sum_cd<NR>(t2);
      if (f.zero) t1 = t2 = mk(NAN, NAN);
      if (r == 0) {
        rk[1] = t1;
        rk[2] = t2;
      }
    }
  }
}

// The log-sum-exp over the determinants and the Jastrow factor of k_det_rows's results
// res [cfg][K][nt + 2], one wave per configuration (kept out of the determinant kernel, whose
// register budget it would otherwise set).
template <int KIND>
__global__ __launch_bounds__(256) void k_det_tail(JetGeom j, DetParams p, const double* __restrict__ R,
                                                  const int32_t* __restrict__ S, int64_t w0,
                                                  const cd* __restrict__ res, double* __restrict__ out_lp,
                                                  double* __restrict__ out_tan) {
  __shared__ double pos[4][64];
  __shared__ int spin[4][32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t cfg = (int64_t)blockIdx.x * 4 + wave;
  if (cfg >= j.n_cfg) return;  // the whole wave; no workgroup barrier follows
  const int N = j.N, rs_ = j.nt + 2;
  int64_t w;
  int flip = -1;
  if (j.mode == 2) {
    w = w0 + cfg / N;
    flip = (int)(cfg % N);
  } else {
    w = w0 + cfg;
  }
  if (p.use_jastrow && lane < N) {
    pos[wave][2 * lane] = R[(w * N + lane) * 2];
    pos[wave][2 * lane + 1] = R[(w * N + lane) * 2 + 1];
    const int s = S[w * N + lane];
    spin[wave][lane] = lane == flip ? 1 - s : s;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  det_tail_lanes<KIND>(j, p, res + cfg * p.K * rs_, rs_, pos[wave], spin[wave], cfg, flip, out_lp, out_tan);
}

// ------------------------------------------------------------------ wave-level head for 16 < N <= 32
// The N <= 16 scheme on 32 x 32 matrices (padded with an identity block): a matrix is 2 x 2
// blocks of 16 x 16, each in one of the two layouts above, lane = 16 g + l:
//   row chunk    slot [2 bi + bj][s] holds X[16 bi + l][16 bj + 4 g + s]   (M, LU factors, Mi)
//   column chunk slot [2 bi + bj][s] holds X[16 bj + 4 g + s][16 bi + l]   (dM)
// so tr(Mi dM) is a lane-local dot product over the 16 slots, and every 16 x 16 block of
// Mi dM is eight v_mfma_f64_16x16x4f64 steps (two inner blocks of four).  64 registers per
// matrix; rows move with cross-lane shuffles as in the N <= 16 head.
struct Lu32 {
  cd m[4][4];   // row chunk; after factor(): L below the diagonal (unit), U on and above
  int orig[2];  // original index of the row now at position 16 bi + l
  int parity;
  bool zero_pivot;
};

// Pivot step q = 16 BQ + QL at compile time (column q in block column BQ, slot QL & 3 of lanes
// 16 (QL >> 2) + l).  With BQ = 1 rows 0..15 are final and the pivot row lies in block row 1, so
// every block index is static.  Broadcasts of row q and the pivot search stay inside the 16-lane
// rows (DPP); the column and the rows of the (run-time) pivot move with ds_bpermute.
template <int BQ, int QL>
__device__ __forceinline__ void lu32_step(Lu32& f, int lane) {
  constexpr int q = 16 * BQ + QL, gq = QL >> 2, sq = QL & 3, lq = QL;
  const int l = lane & 15, g = lane >> 4;
  // column q of this lane's rows l (block row 0) and 16 + l (block row 1)
  cd c0 = BQ == 0 ? shfl_cd(f.m[0][sq], l + 16 * gq) : mk(0, 0);
  cd c1 = shfl_cd(f.m[2 + BQ][sq], l + 16 * gq);
  const double v0 = (BQ == 0 && l >= q) ? fabs(c0.re) + fabs(c0.im) : -1.0;
  const double v1 = 16 + l >= q ? fabs(c1.re) + fabs(c1.im) : -1.0;
  double v = v1 > v0 ? v1 : v0;
  int br = v1 > v0 ? 16 + l : l;
#define NQS_LU_MAX(K)                                                        \
  {                                                                          \
    const double ov = dpp_d<0x120 + K>(v);                                   \
    const int obr = dpp_i<0x120 + K>(br);                                    \
    const bool take = ov > v || (ov == v && obr < br);                       \
    v = take ? ov : v;                                                       \
    br = take ? obr : br;                                                    \
  }
  NQS_LU_MAX(1)
  NQS_LU_MAX(2)
  NQS_LU_MAX(4)
  NQS_LU_MAX(8)
#undef NQS_LU_MAX
  const int p = __builtin_amdgcn_readfirstlane(br);
  f.zero_pivot = f.zero_pivot || __builtin_amdgcn_readfirstlane(v == 0.0 ? 1 : 0) != 0;
  const int bp = BQ == 1 ? 1 : p >> 4, lp = p & 15;
  // swap rows p and q (branch-free: with p == q every write stores the value it replaces, and
  // the matrix keeps one set of registers)
  f.parity ^= p != q ? 1 : 0;
#pragma unroll
  for (int bj = 0; bj < 2; ++bj)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      cd& mq = f.m[2 * BQ + bj][s];
      const cd xp = shfl_cd(bp ? f.m[2 + bj][s] : f.m[bj][s], lp + 16 * g);
      const cd xq = bcast16<lq>(mq);
      if (l == lq) mq = xp;
      if (l == lp) {
        if (bp) f.m[2 + bj][s] = xq;
        else f.m[bj][s] = xq;
      }
    }
  const int op = __shfl(bp ? f.orig[1] : f.orig[0], lp + 16 * g);
  const int oq = dpp_i<0x150 + lq>(f.orig[BQ]);
  if (l == lq) f.orig[BQ] = op;
  if (l == lp) {
    if (bp) f.orig[1] = oq;
    else f.orig[0] = oq;
  }
  // c0 / c1 are the same in every row of 16 lanes: the pivot's and row q's values by readlane
  const cd cp = bp ? c1 : c0, cqq = BQ ? c1 : c0;
  const cd xpq = mk(readlane_d(cp.re, lp), readlane_d(cp.im, lp));
  const cd xqq = mk(readlane_d(cqq.re, lq), readlane_d(cqq.im, lq));
  if (l == lq) {
    if (BQ) c1 = xpq;
    else c0 = xpq;
  }
  if (l == lp) {
    if (bp) c1 = xqq;
    else c0 = xqq;
  }
  const cd pv = xpq;
  const bool zero = pv.re == 0.0 && pv.im == 0.0;
  const cd l0 = (BQ == 0 && l > q && !zero) ? cdiv_(c0, pv) : mk(0, 0);
  const cd l1 = (16 + l > q && !zero) ? cdiv_(c1, pv) : mk(0, 0);
  // eliminate below row q (columns of block column 0 are final once q >= 16)
#pragma unroll
  for (int bj = BQ; bj < 2; ++bj)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const cd rq = bcast16<lq>(f.m[2 * BQ + bj][s]);
      const int col = 16 * bj + 4 * g + s;
      if (BQ == 0 && l > q) {
        if (col == q) f.m[bj][s] = l0;
        else if (col > q) f.m[bj][s] = f.m[bj][s] - l0 * rq;
      }
      if (16 + l > q) {
        if (col == q) f.m[2 + bj][s] = l1;
        else if (col > q) f.m[2 + bj][s] = f.m[2 + bj][s] - l1 * rq;
      }
    }
}

template <int BQ, int QL>
struct Lu32Steps {
  static __device__ __forceinline__ void run(Lu32& f, int lane) {
    lu32_step<BQ, QL>(f, lane);
    Lu32Steps<BQ, QL + 1>::run(f, lane);
  }
};
template <int BQ>
struct Lu32Steps<BQ, 16> {
  static __device__ __forceinline__ void run(Lu32&, int) {}
};

__device__ __forceinline__ void lu32_factor(Lu32& f, int lane) {
  const int l = lane & 15;
  f.orig[0] = l;
  f.orig[1] = 16 + l;
  f.parity = 0;
  f.zero_pivot = false;
  Lu32Steps<0, 0>::run(f, lane);
  Lu32Steps<1, 0>::run(f, lane);
}

// log det from the factors (principal branch), as lu_logdet for the first N pivots
__device__ __forceinline__ cd lu32_logdet(const Lu32& f, int lane, int N) {
  const int l = lane & 15, g = lane >> 4;
  const bool on = (l >> 2) == g;  // this lane holds U[l][l] (block 0) and U[16 + l][16 + l] (block 3)
  const cd u0 = pick4(f.m[0], l & 3), u1 = pick4(f.m[3], l & 3);
  const bool m0 = on && l < N, m1 = on && 16 + l < N;
  const double a0 = m0 ? cabs_(u0) : 1.0, a1 = m1 ? cabs_(u1) : 1.0;
  double la = wave_sum_d(log(a0) + log(a1));
  cd ph0 = m0 && a0 > 0 ? mk(u0.re / a0, u0.im / a0) : mk(1, 0);
  cd ph1 = m1 && a1 > 0 ? mk(u1.re / a1, u1.im / a1) : mk(1, 0);
  cd ph = wave_prod_cd(ph0 * ph1);
  if (f.parity) ph = mk(-ph.re, -ph.im);
  return mk(la + log(hypot(ph.re, ph.im)), atan2(ph.im, ph.re));
}

// Mi = U^-1 L^-1 P in the row-chunk layout.  Forward: rows below q minus L[r][q] row q;
// backward: row q divided by U[q][q], rows above q minus U[r][q] row q.
// keeps x in registers at this point: the updates of rows that no later step of a half reads are
// not sunk to its end (which would hold every step's broadcast row in registers)
__device__ __forceinline__ void pin_cd(cd& x) { asm volatile("" : "+v"(x.re), "+v"(x.im)); }

template <int BQ, int QL>
struct Lu32Inv {
  static constexpr int q = 16 * BQ + QL, gq = QL >> 2, sq = QL & 3, lq = QL;
  static __device__ __forceinline__ void fwd(const Lu32& f, int lane, cd (&y)[4][4]) {
    const int l = lane & 15;
    const cd c0 = BQ == 0 ? shfl_cd(f.m[0][sq], l + 16 * gq) : mk(0, 0);
    const cd c1 = shfl_cd(f.m[2 + BQ][sq], l + 16 * gq);
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const cd yq = bcast16<lq>(y[2 * BQ + bj][s]);
        if (BQ == 0 && l > q) y[bj][s] = y[bj][s] - c0 * yq;
        if (16 + l > q) y[2 + bj][s] = y[2 + bj][s] - c1 * yq;
        pin_cd(y[2 + bj][s]);
      }
    Lu32Inv<BQ, QL + 1>::fwd(f, lane, y);
  }
  static __device__ __forceinline__ void bwd(const Lu32& f, int lane, cd (&y)[4][4]) {
    const int l = lane & 15;
    constexpr int UL = lq + 16 * gq;  // lane of U[q][q]
    const cd uqq = f.m[3 * BQ][sq];
    const cd ru = cdiv_(mk(1, 0), mk(readlane_d(uqq.re, UL), readlane_d(uqq.im, UL)));  // 1 / U[q][q]
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        cd& yy = y[2 * BQ + bj][s];
        const cd dv = yy * ru;
        if (l == lq) yy = dv;
      }
    const cd c0 = shfl_cd(f.m[BQ][sq], l + 16 * gq);
    const cd c1 = BQ == 1 ? shfl_cd(f.m[3][sq], l + 16 * gq) : mk(0, 0);
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const cd yq = bcast16<lq>(y[2 * BQ + bj][s]);
        if (l < q) y[bj][s] = y[bj][s] - c0 * yq;
        if (BQ == 1 && 16 + l < q) y[2 + bj][s] = y[2 + bj][s] - c1 * yq;
        pin_cd(y[bj][s]);
      }
    Lu32Inv<BQ, QL - 1>::bwd(f, lane, y);
  }
};
template <int BQ>
struct Lu32Inv<BQ, 16> {
  static __device__ __forceinline__ void fwd(const Lu32&, int, cd (&)[4][4]) {}
};
template <int BQ>
struct Lu32Inv<BQ, -1> {
  static __device__ __forceinline__ void bwd(const Lu32&, int, cd (&)[4][4]) {}
};

__device__ __forceinline__ void lu32_inverse(const Lu32& f, int lane, cd (&y)[4][4]) {
  const int g = lane >> 4;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int s = 0; s < 4; ++s) y[2 * bi + bj][s] = mk(16 * bj + 4 * g + s == f.orig[bi] ? 1.0 : 0.0, 0);
  Lu32Inv<0, 0>::fwd(f, lane, y);
  Lu32Inv<1, 0>::fwd(f, lane, y);
  Lu32Inv<1, 15>::bwd(f, lane, y);
  Lu32Inv<0, 15>::bwd(f, lane, y);
}

// sum over the wave of tr(X Y) for X = A_IJ, Y = the accumulators of (A_JI)^T (f64 MFMA C
// layouts of the same 16 x 16 block): one complex product per accumulator slot
__device__ __forceinline__ cd acc_dot(const d4v& are, const d4v& aim, const d4v& tre, const d4v& tim) {
  cd h = mk(0, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r) h = h + mk(are[r], aim[r]) * mk(tre[r], tim[r]);
  return h;
}

// block (I, J) of A = Mi dM and of A^T's block (I, J) = (A_JI)^T: Mi row chunk (A operand of
// the MFMA, k index 16 B + 4 g + s), dM column chunk (B operand); A^T from the same registers
// with the operands swapped
__device__ __forceinline__ void mfma_block_pair(const cd (&mi)[4][4], const cd (&dm)[4][4], int I, int J, d4v& are,
                                                d4v& aim, d4v& tre, d4v& tim) {
  are = (d4v)0.0;
  aim = (d4v)0.0;
  tre = (d4v)0.0;
  tim = (d4v)0.0;
#pragma unroll
  for (int B = 0; B < 2; ++B)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const cd x = mi[2 * I + B][s], y = dm[2 * J + B][s];  // A_IJ += Mi_IB dM_BJ
      are = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re, y.re, are, 0, 0, 0);
      are = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.im, y.im, are, 0, 0, 0);
      aim = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re, y.im, aim, 0, 0, 0);
      aim = __builtin_amdgcn_mfma_f64_16x16x4f64(x.im, y.re, aim, 0, 0, 0);
      const cd u = dm[2 * I + B][s], v = mi[2 * J + B][s];  // (A_JI)^T += (dM_BI)^T (Mi_JB)^T
      tre = __builtin_amdgcn_mfma_f64_16x16x4f64(u.re, v.re, tre, 0, 0, 0);
      tre = __builtin_amdgcn_mfma_f64_16x16x4f64(-u.im, v.im, tre, 0, 0, 0);
      tim = __builtin_amdgcn_mfma_f64_16x16x4f64(u.re, v.im, tim, 0, 0, 0);
      tim = __builtin_amdgcn_mfma_f64_16x16x4f64(u.im, v.re, tim, 0, 0, 0);
    }
}

// One workgroup per configuration (1, 2 or 4 waves); envelope jet built once in LDS, then each
// wave takes determinants k = wave, wave + nwaves, ... in registers.  KIND as k_det_jet16.
template <typename T, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KIND == 2 ? 1 : 2)))
void k_det_jet32(JetGeom j, DetParams p, const double* __restrict__ R,
                                                   const int32_t* __restrict__ S, int64_t w0,
                                                   const T* __restrict__ bf, double* __restrict__ out_lp,
                                                   double* __restrict__ out_tan, double* __restrict__ out_lap) {
  // LDS: env [32 x 32] (entry (a, i) at 32 a + i), then dex, dey (KIND >= 1), lenv (KIND 2); then
  // a region holding the phase table exp(i G_g . r_i) [g][i] while the jet is built and res
  // [K][nt + 2] (log det, tangents, Laplacian term per determinant) afterwards
  extern __shared__ cd sh32[];
  constexpr int NJ = KIND == 0 ? 1 : (KIND == 1 ? 3 : 4);
  cd* env = sh32;
  cd* dex = sh32 + 1024;
  cd* dey = sh32 + 2048;
  cd* lenv = sh32 + 3072;
  cd* tail = sh32 + NJ * 1024;
  __shared__ double pos[64];
  __shared__ int spin[32];
  const int N = j.N, K = p.K, nt = j.nt;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  const int64_t cfg = blockIdx.x;
  int64_t w;
  int flip = -1;
  if (j.mode == 2) {
    w = w0 + cfg / N;
    flip = (int)(cfg % N);
  } else {
    w = w0 + cfg;
  }
  if (tid < N) {
    pos[2 * tid] = R[(w * N + tid) * 2];
    pos[2 * tid + 1] = R[(w * N + tid) * 2 + 1];
    const int s = S[w * N + tid];
    spin[tid] = (tid == flip) ? 1 - s : s;
  }
  __syncthreads();
  const bool tab = p.n_g <= kPhaseMax;
  if (tab) {
    for (int e = tid; e < p.n_g * 32; e += blockDim.x) {
      const int gg = e >> 5, i = e & 31;
      double sn = 0, cs = 0;
      if (i < N) sincos(p.gs[2 * gg] * pos[2 * i] + p.gs[2 * gg + 1] * pos[2 * i + 1], &sn, &cs);
      tail[e] = mk(cs, sn);
    }
    __syncthreads();
  }
  for (int e = tid; e < 1024; e += blockDim.x) {  // envelope jet, entry (a, i) = (e / 32, e % 32)
    const int a = e >> 5, i = e & 31;
    cd v = mk(0, 0), gx = mk(0, 0), gy = mk(0, 0), lp = mk(0, 0);
    if (a < N && i < N) {
      for (int gg = 0; gg < p.n_g; ++gg) {
        const double Gx = p.gs[2 * gg], Gy = p.gs[2 * gg + 1];
        double sn, cs;
        if (tab) {
          cs = tail[gg * 32 + i].re;
          sn = tail[gg * 32 + i].im;
        } else {
          sincos(Gx * pos[2 * i] + Gy * pos[2 * i + 1], &sn, &cs);
        }
        const double* wp = p.wenv + ((size_t)(a * p.n_g + gg) * 2 + spin[i]) * 2;
        const cd t = mk(wp[0], wp[1]) * mk(cs, sn);
        v = v + t;
        if constexpr (KIND > 0) {
          gx = gx + t * mk(0, Gx);
          gy = gy + t * mk(0, Gy);
        }
        if constexpr (KIND == 2) lp = lp + (-(Gx * Gx + Gy * Gy)) * t;
      }
    }
    env[e] = v;
    if constexpr (KIND > 0) {
      dex[e] = gx;
      dey[e] = gy;
    }
    if constexpr (KIND == 2) lenv[e] = lp;
  }
  __syncthreads();  // the phase table is dead from here; tail holds res
  cd* res = tail;
  const int C = j.C;
  const int64_t KN = (int64_t)K * N, rowlen = 2 * KN, chs = (int64_t)N * rowlen;
  const T* bfc0 = bf + cfg * C * chs;
  const int l = lane & 15, g = lane >> 4;
  const int rs_ = nt + 2;
  for (int k = wave; k < K; k += nwaves) {
    Lu32 f;
    {
      const int l = opaque_i(lane & 15), g = opaque_i(lane >> 4);
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int a = 16 * bi + l, i = 16 * bj + 4 * g + s;
            const cd b = bf_at(bfc0, rowlen, KN, N, k, a, i);
            f.m[2 * bi + bj][s] = (a < N && i < N) ? env[a * 32 + i] * (mk(1, 0) + b) : mk(a == i ? 1.0 : 0.0, 0);
            __builtin_amdgcn_sched_barrier(0);
          }
    }
    lu32_factor(f, lane);
    const cd ld = lu32_logdet(f, lane, N);
    if (lane == 0) res[k * rs_] = ld;
    if constexpr (KIND == 0) continue;
    cd mi[4][4];
    lu32_inverse(f, lane, mi);
    cd lap_tr = mk(0, 0);
    if constexpr (KIND == 2) {  // tr(Mi Lap M) first, while only Mi is held in registers
      const int cl = nt + 1;
      const T* bfl = bfc0 + cl * chs;
      const int l = opaque_i(lane & 15), g = opaque_i(lane >> 4);
      cd tr = mk(0, 0);
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int a = 16 * bj + 4 * g + s, i = 16 * bi + l, e_ = a * 32 + i;
            const cd ob = mk(1, 0) + bf_at(bfc0, rowlen, KN, N, k, a, i);
            const cd v = lenv[e_] * ob + env[e_] * bf_at(bfl, rowlen, KN, N, k, a, i);
            tr = tr + mi[2 * bi + bj][s] * v;
            __builtin_amdgcn_sched_barrier(0);
          }
      tr = wave_sum_cd(tr);
      lap_tr = tr;  // plus the cross terms and minus sum_c tr((Mi dM_c)^2), below
    }
    // Lap M = Lap env (1 + bf) + env Lap bf + 2 sum_c denv_c dbf_c: the cross term's column i is
    // tangent c = 1 + 2 i + d's own column, so it is summed in the tangent loop
    cd lapx = mk(0, 0);
    cd hsum = mk(0, 0);
    for (int c = 1; c <= nt; ++c) {
      int ec, dc;
      if (j.mode == 2) {
        ec = flip;
        dc = c - 1;
      } else {
        ec = (c - 1) >> 1;
        dc = (c - 1) & 1;
      }
      const T* bfc = bfc0 + c * chs;
      const cd* den = dc == 0 ? dex : dey;
      const int l = opaque_i(lane & 15), g = opaque_i(lane >> 4);
      cd dm[4][4];
      cd tr = mk(0, 0);
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int a = 16 * bj + 4 * g + s, i = 16 * bi + l, e_ = a * 32 + i;
            const cd bv = bf_at(bfc, rowlen, KN, N, k, a, i);
            cd v = env[e_] * bv;
            if (i == ec && a < N) {
              const cd de = den[e_];
              v = v + de * (mk(1, 0) + bf_at(bfc0, rowlen, KN, N, k, a, i));
              if constexpr (KIND == 2) lapx = lapx + mi[2 * bi + bj][s] * (2.0 * (de * bv));
            }
            dm[2 * bi + bj][s] = v;
            tr = tr + mi[2 * bi + bj][s] * v;
            __builtin_amdgcn_sched_barrier(0);
          }
      tr = wave_sum_cd(tr);
      if (lane == 0) res[k * rs_ + c] = tr;
      if constexpr (KIND == 2) {
        // tr(A^2) = tr(A_00^2) + tr(A_11^2) + 2 tr(A_01 A_10)
        d4v are, aim, tre, tim;
        mfma_block_pair(mi, dm, 0, 0, are, aim, tre, tim);
        cd h = acc_dot(are, aim, tre, tim);
        mfma_block_pair(mi, dm, 1, 1, are, aim, tre, tim);
        h = h + acc_dot(are, aim, tre, tim);
        mfma_block_pair(mi, dm, 0, 1, are, aim, tre, tim);
        h = h + 2.0 * acc_dot(are, aim, tre, tim);
        hsum = hsum + wave_sum_cd(h);
      }
    }
    if constexpr (KIND == 2) {
      lap_tr = lap_tr + wave_sum_cd(lapx);
      if (lane == 0) res[k * rs_ + nt + 1] = lap_tr - hsum;
    }
  }
  __syncthreads();
  if (wave != 0) return;
  det_head_tail<KIND>(j, p, res, rs_, pos, spin, cfg, flip, out_lp, out_tan, out_lap);
}

// Note on rint vs jnp.round: both round half to even; minimum image is exact either way.
template <typename T>
int launch_det_jet(const JetGeom& j, const DetParams& p, const double* R, const int32_t* S, int64_t w0,
                   const T* bf, double* out_lp, double* out_tan, double* out_lap, void* scratch, size_t scratch_bytes,
                   hipStream_t st) {
  if (j.n_cfg == 0) return NQS_OK;
  const int N = j.N, NN = N * N;
  // value-only and spin-flip passes: one determinant per group of NR lanes (NR = 16 for N <= 16,
  // 32 up to N = 32), one row per lane, then the log-sum-exp per configuration; the
  // per-determinant results are staged in `scratch`
  if (N <= 32 && p.K <= 16 && !j.has_lap && (j.nt == 0 || (j.mode == 2 && j.nt == 2))) {
    const int kind = j.nt == 0 ? 0 : 1;
    const int NR = N <= 16 ? 16 : 32, GW = 64 / NR;
    int G = std::max(1, std::min(4, 16 / p.K));
    while (G > 1 && rowdet_lds(NR, G, N, p.n_g, kind) > 64 * 1024) --G;
    const int waves = std::min(16 / GW, (G * p.K + GW - 1) / GW);
    const size_t lds = rowdet_lds(NR, G, N, p.n_g, kind);
    const size_t need = (size_t)j.n_cfg * p.K * (j.nt + 2) * sizeof(cd);
    if (lds <= 64 * 1024 && scratch != nullptr && need <= scratch_bytes) {
      cd* res = static_cast<cd*>(scratch);
      const dim3 grid((unsigned)cdiv(j.n_cfg, G)), block(64 * waves), tgrid((unsigned)cdiv(j.n_cfg, 4));
      if (NR == 16) {
        if (kind == 0)
          hipLaunchKernelGGL((k_det_rows<T, 0, 16>), grid, block, lds, st, j, p, G, R, S, w0, bf, res);
        else
          hipLaunchKernelGGL((k_det_rows<T, 1, 16>), grid, block, lds, st, j, p, G, R, S, w0, bf, res);
      } else {
        if (kind == 0)
          hipLaunchKernelGGL((k_det_rows<T, 0, 32>), grid, block, lds, st, j, p, G, R, S, w0, bf, res);
        else
          hipLaunchKernelGGL((k_det_rows<T, 1, 32>), grid, block, lds, st, j, p, G, R, S, w0, bf, res);
      }
      NQS_LAUNCH_CHECK();
      if (kind == 0)
        hipLaunchKernelGGL(k_det_tail<0>, tgrid, dim3(256), 0, st, j, p, R, S, w0, res, out_lp, out_tan);
      else
        hipLaunchKernelGGL(k_det_tail<1>, tgrid, dim3(256), 0, st, j, p, R, S, w0, res, out_lp, out_tan);
      NQS_LAUNCH_CHECK();
      return NQS_OK;
    }
  }
  if (N <= 16 && j.nt <= 64) {
    const size_t lds = sizeof(cd) * (size_t)p.K * (j.nt + 2);
    NQS_REQUIRE(lds <= 96 * 1024, NQS_ERR_UNSUPPORTED, "determinant head exceeds LDS");
    // workgroup size (determinants in flight per configuration) chosen by the autotuner
    static const char* const names[3] = {"2waves", "4waves", "8waves"};
    const TuneKey key{TUNE_DET16, (int64_t)sizeof(T) * 1000 + N * 64 + p.K, (int64_t)j.nt * 4 + j.has_lap * 2 +
                      (j.mode == 2), tune_bucket(j.n_cfg)};
    const int kind = j.has_lap ? 2 : (j.nt > 0 ? 1 : 0);
    return tune_run(key, 3, 1, names, [&](int c) -> int {
      const dim3 grid((unsigned)j.n_cfg), block(128 << c);
      if (kind == 0)
        hipLaunchKernelGGL((k_det_jet16<T, 0>), grid, block, lds, st, j, p, R, S, w0, bf, out_lp, out_tan, out_lap);
      else if (kind == 1)
        hipLaunchKernelGGL((k_det_jet16<T, 1>), grid, block, lds, st, j, p, R, S, w0, bf, out_lp, out_tan, out_lap);
      else
        hipLaunchKernelGGL((k_det_jet16<T, 2>), grid, block, lds, st, j, p, R, S, w0, bf, out_lp, out_tan, out_lap);
      NQS_LAUNCH_CHECK();
      return NQS_OK;
    }, st);
  }
  NQS_REQUIRE(NN <= 4 * 256, NQS_ERR_UNSUPPORTED, "determinant head supports N <= 32");
  if (j.nt <= 64) {
    const int kind = j.has_lap ? 2 : (j.nt > 0 ? 1 : 0);
    const int nj = kind == 0 ? 1 : (kind == 1 ? 3 : 4);
    const int64_t tab = p.n_g <= kPhaseMax ? (int64_t)p.n_g * 32 : 0;
    const int64_t rs = (int64_t)p.K * (j.nt + 2);
    const size_t lds = sizeof(cd) * (size_t)(nj * 1024 + (tab > rs ? tab : rs));
    if (lds + 1024 <= 160 * 1024) {  // + the static positions and spins
      static size_t lds_set[3] = {0, 0, 0};
      if (lds > 64 * 1024 && lds > lds_set[kind]) {
        const void* fn = kind == 0 ? (const void*)k_det_jet32<T, 0>
                                   : (kind == 1 ? (const void*)k_det_jet32<T, 1> : (const void*)k_det_jet32<T, 2>);
        NQS_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        lds_set[kind] = lds;
      }
      // determinants in flight per configuration (waves per workgroup) chosen by the autotuner
      static const char* const names[3] = {"1wave", "2waves", "4waves"};
      const TuneKey key{TUNE_DET32, (int64_t)sizeof(T) * 1000 + N * 64 + p.K,
                        (int64_t)j.nt * 4 + j.has_lap * 2 + (j.mode == 2), tune_bucket(j.n_cfg)};
      return tune_run(key, 3, 2, names, [&](int c) -> int {
        const dim3 grid((unsigned)j.n_cfg), block(64 << c);
        if (kind == 0)
          hipLaunchKernelGGL((k_det_jet32<T, 0>), grid, block, lds, st, j, p, R, S, w0, bf, out_lp, out_tan, out_lap);
        else if (kind == 1)
          hipLaunchKernelGGL((k_det_jet32<T, 1>), grid, block, lds, st, j, p, R, S, w0, bf, out_lp, out_tan, out_lap);
        else
          hipLaunchKernelGGL((k_det_jet32<T, 2>), grid, block, lds, st, j, p, R, S, w0, bf, out_lp, out_tan, out_lap);
        NQS_LAUNCH_CHECK();
        return NQS_OK;
      }, st);
    }
  }
  size_t lds = sizeof(cd) * (size_t)(7 * NN + (size_t)p.K * (j.nt + 2)) + sizeof(double) * 2 * N +
               sizeof(int) * (2 * N + 4) + sizeof(double) * (size_t)(8 + 2 + 2 * (j.nt + 1) + (j.nt + 1) + 8);
  lds = round_up((int64_t)lds, 16);
  NQS_REQUIRE(lds <= 160 * 1024, NQS_ERR_UNSUPPORTED, "determinant head exceeds LDS");
  hipLaunchKernelGGL(k_det_jet<T>, dim3((unsigned)j.n_cfg), dim3(256), lds, st, j, p, R, S, w0, bf, out_lp,
                     out_tan, out_lap);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

template int launch_det_jet<double>(const JetGeom&, const DetParams&, const double*, const int32_t*, int64_t,
                                    const double*, double*, double*, double*, void*, size_t, hipStream_t);
template int launch_det_jet<float>(const JetGeom&, const DetParams&, const double*, const int32_t*, int64_t,
                                   const float*, double*, double*, double*, void*, size_t, hipStream_t);


// ------------------------------------------------------------------ determinant-head backward
// dL/dM_k[a,i] = p_k Minv_k[i,a] (p_k the log-sum-exp weights, 1 for K = 1), hence
//   dL/dbf[k,a,i] = p_k env[a,i] Minv_k[i,a],   dL/denv[a,i] = sum_k p_k (1 + bf[k,a,i]) Minv_k[i,a].
// The backflow Dense has 2KN real outputs y = [Re bf | Im bf]; its complex cotangent is
// [dL/dbf, i dL/dbf], stored as two real channels (Re, Im).  Complex leaves (det_logw, dW)
// give O(re part) = dL/dtheta and O(im part) = i dL/dtheta.
// O entry q (< 2 nk, nk = n_knots for the RBF Jastrow, 1 for the cusp) of the Jastrow leaves
// for one configuration (pos [N][2], spin [N]): q < nk the same-spin knot q (the cusp's loga),
// else the opposite-spin knot q - nk.  Returns the entry's theta offset, -1 when it has no leaf
// (the cusp has no opposite-spin entry).
__device__ __forceinline__ int64_t jastrow_o_entry(const DetParams& p, const DetBwdOffsets& off, int N,
                                                   const double* pos, const int* spin, int q, double& v) {
  const int nk = p.jastrow_kind == 0 ? p.n_knots : 1;
  const bool same_leaf = q < nk;
  const int kk = q % nk;
  v = 0;
  if (p.jastrow_kind != 0 && !same_leaf) return -1;
  for (int a = 0; a < N; ++a)
    for (int b = a + 1; b < N; ++b) {
      const bool same = spin[a] == spin[b];
      const double dx = pos[2 * a] - pos[2 * b], dy = pos[2 * a + 1] - pos[2 * b + 1];
      double r, grx, gry, lapr;
      pair_dist(p, dx, dy, r, grx, gry, lapr);
      if (p.jastrow_kind == 0) {
        if (same != same_leaf) continue;
        const double mu = (p.n_knots > 1) ? p.rbf_rcut * (double)kk / (double)(p.n_knots - 1) : 0.0;
        const double z = (r - mu) / p.rbf_w;
        v += exp(-z * z);
      } else {
        const double la = p.jas[0];
        const double aa = (la > 0 ? la + log1p(exp(-la)) : log1p(exp(la))) + 0.05;
        const double sig = 1.0 / (1.0 + exp(-la));
        const double al = same ? p.cusp / 3.0 : p.cusp;
        const double e = exp(-r / aa);
        v += (al * (1.0 - e) - al * (r / aa) * e) * sig;
      }
    }
  if (p.jastrow_kind == 0) return same_leaf ? off.jas_same + kk : off.jas_opp + kk;
  return off.jas_loga;
}

template <typename T>
__global__ __launch_bounds__(256) void k_det_bwd(DetParams p, const double* __restrict__ R,
                                                 const int32_t* __restrict__ S, int64_t w0, const T* __restrict__ bf,
                                                 T* __restrict__ ybar, double* __restrict__ out_lp, T* __restrict__ O,
                                                 int64_t ldo, int64_t W_total, DetBwdOffsets off) {
  extern __shared__ double shd[];
  const int N = p.N, NN = N * N, K = p.K;
  const int tid = threadIdx.x, nthr = blockDim.x;
  cd* env = reinterpret_cast<cd*>(shd);
  cd* envb = env + NN;
  cd* M = envb + NN;
  cd* Mi = M + NN;
  cd* res = Mi + NN;  // K: logdet, then weights
  double* pos = reinterpret_cast<double*>(res + K);
  int* spin = reinterpret_cast<int*>(pos + 2 * N);
  int* piv = spin + N;
  int* flag = piv + N;
  const int64_t cfg = blockIdx.x;
  const int64_t w = w0 + cfg;
  for (int i = tid; i < N; i += nthr) {
    pos[2 * i] = R[(w * N + i) * 2];
    pos[2 * i + 1] = R[(w * N + i) * 2 + 1];
    spin[i] = S[w * N + i];
  }
  __syncthreads();
  for (int e = tid; e < NN; e += nthr) {
    const int a = e / N, i = e % N;
    cd v = mk(0, 0);
    for (int g = 0; g < p.n_g; ++g) {
      double sn, cs;
      sincos(p.gs[2 * g] * pos[2 * i] + p.gs[2 * g + 1] * pos[2 * i + 1], &sn, &cs);
      const double* wp = p.wenv + ((size_t)(a * p.n_g + g) * 2 + spin[i]) * 2;
      v = v + mk(wp[0], wp[1]) * mk(cs, sn);
    }
    env[e] = v;
    envb[e] = mk(0, 0);
  }
  __syncthreads();
  const int64_t KN = (int64_t)K * N, rowlen = 2 * KN;
  const T* bfc = bf + cfg * (int64_t)N * rowlen;
  auto bfval = [&](int k, int a, int i) -> cd {
    const T* r = bfc + (int64_t)i * rowlen;
    return mk((double)r[k * N + a], (double)r[KN + k * N + a]);
  };
  // pass 1: log dets
  for (int k = 0; k < K; ++k) {
    for (int e = tid; e < NN; e += nthr) M[e] = env[e] * (mk(1, 0) + bfval(k, e / N, e % N));
    __syncthreads();
    wg_lu(M, N, piv, flag);
    if (tid == 0) res[k] = lu_logdet(M, N, flag[0]);
    __syncthreads();
  }
  if (tid == 0) {
    cd L;
    if (K == 1) {
      L = res[0];
      res[0] = mk(1, 0);
    } else {
      double amax = -INFINITY;
      for (int k = 0; k < K; ++k) {
        res[k] = res[k] + mk(p.det_logw[2 * k], p.det_logw[2 * k + 1]);
        amax = res[k].re > amax ? res[k].re : amax;
      }
      if (!isfinite(amax)) amax = 0.0;
      cd s = mk(0, 0);
      for (int k = 0; k < K; ++k) s = s + cexp_(res[k] - mk(amax, 0));
      L = clog_(s) + mk(amax, 0);
      for (int k = 0; k < K; ++k) res[k] = cexp_(res[k] - L);
    }
    double J = 0;
    if (p.use_jastrow) {
      for (int a = 0; a < N; ++a)
        for (int b = a + 1; b < N; ++b) {
          const double dx = pos[2 * a] - pos[2 * b], dy = pos[2 * a + 1] - pos[2 * b + 1];
          double r, grx, gry, lapr;
          pair_dist(p, dx, dy, r, grx, gry, lapr);
          double u, u1, u2;
          jastrow_u(p, spin[a] == spin[b], r, u, u1, u2);
          J += u;
        }
    }
    if (out_lp) {
      out_lp[2 * cfg] = L.re + J;
      out_lp[2 * cfg + 1] = L.im;
    }
  }
  __syncthreads();
  // pass 2: inverses, backflow cotangent, envelope cotangent
  T* y0 = ybar + (cfg * 2 + 0) * (int64_t)N * rowlen;
  T* y1 = ybar + (cfg * 2 + 1) * (int64_t)N * rowlen;
  for (int k = 0; k < K; ++k) {
    for (int e = tid; e < NN; e += nthr) M[e] = env[e] * (mk(1, 0) + bfval(k, e / N, e % N));
    __syncthreads();
    wg_lu(M, N, piv, flag);
    wg_inverse(M, Mi, N, piv);
    const cd pk = res[k];
    for (int e = tid; e < NN; e += nthr) {
      const int a = e / N, i = e % N;
      const cd mi = pk * Mi[i * N + a];
      const cd g = env[e] * mi;  // dL/dbf[k,a,i]
      const int64_t col = (int64_t)i * rowlen + k * N + a;
      y0[col] = (T)g.re;
      y0[col + KN] = (T)(-g.im);
      y1[col] = (T)g.im;
      y1[col + KN] = (T)g.re;
      envb[e] = envb[e] + (mk(1, 0) + bfval(k, a, i)) * mi;
    }
    __syncthreads();
  }
  T* Ore = O + (w) * ldo;
  T* Oim = O + (W_total + w) * ldo;
  if (off.dlw >= 0) {
    for (int k = tid; k < K; k += nthr) {
      const cd pk = res[k];
      Ore[off.dlw + k] = (T)pk.re;
      Oim[off.dlw + k] = (T)pk.im;
      Ore[off.dlw + K + k] = (T)(-pk.im);
      Oim[off.dlw + K + k] = (T)pk.re;
    }
  }
  if (off.dW >= 0) {
    const int64_t nW = (int64_t)N * p.n_g * 2;
    for (int64_t q = tid; q < nW; q += nthr) {
      const int a = (int)(q / (p.n_g * 2)), g = (int)((q / 2) % p.n_g), s = (int)(q % 2);
      cd v = mk(0, 0);
      for (int i = 0; i < N; ++i) {
        if (spin[i] != s) continue;
        double sn, cs;
        sincos(p.gs[2 * g] * pos[2 * i] + p.gs[2 * g + 1] * pos[2 * i + 1], &sn, &cs);
        v = v + envb[a * N + i] * mk(cs, sn);
      }
      Ore[off.dW + q] = (T)v.re;
      Oim[off.dW + q] = (T)v.im;
      Ore[off.dW + nW + q] = (T)(-v.im);
      Oim[off.dW + nW + q] = (T)v.re;
    }
  }
  if (p.use_jastrow) {
    const int nk = p.jastrow_kind == 0 ? p.n_knots : 1;
    for (int q = tid; q < 2 * nk; q += nthr) {
      double v;
      const int64_t o = jastrow_o_entry(p, off, N, pos, spin, q, v);
      if (o < 0) continue;
      Ore[o] = (T)v;
      Oim[o] = (T)0;
    }
  }
}

// The same backward, row per lane, for N <= 32 without a learnable envelope (no dW leaf): one
// determinant per group of NR lanes as in k_det_rows (determinant t = (64 / NR) wave + lane / NR of
// the workgroup is k = t % K of configuration t / K), each factored by Gauss-Jordan (KIND 1),
// after which lane r = p_q holds row q of M^-1 with its columns in pivot order, so
// dL/dbf[k,a,q] = p_k env[a,q] M^-1[q,a] is lane-local.  The log dets reach the configuration's
// other determinants through LDS for the log-sum-exp weights p_k.  LDS: rowdet_lds(NR, G, N, n_g,
// 0), then the log dets [G][16].
template <typename T, int NR>
__global__ __launch_bounds__(NR == 16 ? 256 : 512) __attribute__((amdgpu_waves_per_eu(NR == 16 ? 3 : 2)))
void k_det_bwd_rows(JetGeom j, DetParams p, int G, const double* __restrict__ R, const int32_t* __restrict__ S,
                    int64_t w0, const T* __restrict__ bf, T* __restrict__ ybar, double* __restrict__ out_lp,
                    T* __restrict__ O, int64_t ldo, int64_t W_total, DetBwdOffsets off) {
  extern __shared__ __align__(16) char smem[];
  constexpr int EP = NR + 1, GW = 64 / NR;
  const int N = j.N, K = p.K, n_g = p.n_g;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nthr = blockDim.x;
  const int64_t cfg0 = (int64_t)blockIdx.x * G;
  const int ng = (int)min((int64_t)G, j.n_cfg - cfg0);
  cd* env = reinterpret_cast<cd*>(smem);
  double* pos = reinterpret_cast<double*>(env + G * NR * EP);
  int* spin = reinterpret_cast<int*>(pos + G * 2 * NR);
  int* flp = spin + G * NR;
  cd* xrows = reinterpret_cast<cd*>(smem + (rowdet_head(NR, G, 0) + 15) / 16 * 16);
  const bool wlds = N * n_g * 2 * (int)sizeof(cd) <= kRowDetWenvMax;
  cd* wtab = xrows + 16 * NR;
  cd* un = wtab + (wlds ? N * n_g * 2 : 0);
  cd* lds_ld = reinterpret_cast<cd*>(smem + rowdet_lds(NR, G, N, n_g, 0));
  if (wlds)
    for (int e = tid; e < N * n_g * 2; e += nthr) wtab[e] = mk(p.wenv[2 * e], p.wenv[2 * e + 1]);
  const int64_t KN = (int64_t)K * N, rowlen = 2 * KN;
  const int r = lane % NR;
  const int t = GW * wave + lane / NR;
  const int cl = t / K, k = t % K;
  const bool act = cl < ng;  // the same for the NR lanes of a group
  const int64_t cfg = cfg0 + (act ? cl : 0);
  const T* bfc = bf + cfg * (int64_t)N * rowlen;
  // row r of M's backflow factor (NR = 16: loaded before the envelope is built so that its latency
  // overlaps that work)
  T bre[NR], bim[NR];
  auto load_row = [&]() {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const T* q = bfc + (int64_t)(i < N ? i : N - 1) * rowlen + (int64_t)k * N + (r < N ? r : N - 1);
      bre[i] = act ? q[0] : (T)0;
      bim[i] = act ? q[KN] : (T)0;
    }
  };
  if constexpr (NR == 16) load_row();
  rowdet_envelope<0, NR>(j, p, cfg0, ng, R, S, w0, wlds ? wtab : nullptr, un, env, nullptr, pos, spin, flp);
  if constexpr (NR != 16) load_row();
  const cd* envc = env + cl * NR * EP;
  RowDet<NR> f;
  if (act) {
#pragma unroll
    for (int i = 0; i < NR; ++i)
      f.m[i] = (r < N && i < N) ? envc[r * EP + i] * mk(1.0 + (double)bre[i], (double)bim[i])
                                : mk(r == i ? 1.0 : 0.0, 0);
    f.piv = mk(1, 0);
    f.q = r;
#pragma unroll
    for (int w = 0; w < RowPk<NR>::NW; ++w) f.pk[w] = 0;
    f.used = f.zero = false;
    RowDetSteps<0, 1, NR>::run(f, r, xrows + NR * (lane / NR) + 64 * wave, N);
    const cd ld = rowdet_logdet(f, r, N);
    if (r == 0) lds_ld[cl * 16 + k] = ld;
  }
  __syncthreads();
  if (act) {
    // log-sum-exp over the configuration's determinants, in determinant order
    const cd* lz = lds_ld + cl * 16;
    cd L, pk;
    if (K == 1) {
      L = lz[0];
      pk = mk(1, 0);
    } else {
      double amax = -INFINITY;
      for (int q = 0; q < K; ++q) {
        const cd z = lz[q] + mk(p.det_logw[2 * q], p.det_logw[2 * q + 1]);
        amax = z.re > amax ? z.re : amax;
      }
      if (!isfinite(amax)) amax = 0.0;
      cd sum = mk(0, 0);
      for (int q = 0; q < K; ++q) sum = sum + cexp_(lz[q] + mk(p.det_logw[2 * q], p.det_logw[2 * q + 1]) - mk(amax, 0));
      L = clog_(sum) + mk(amax, 0);
      pk = cexp_(lz[k] + mk(p.det_logw[2 * k], p.det_logw[2 * k + 1]) - L);
    }
    const int64_t w = w0 + cfg;
    if (k == 0) {
      double J = 0;
      if (p.use_jastrow) {
        const double* pc = pos + cl * 2 * NR;
        const int* sc = spin + cl * NR;
        double Ja = 0;
        if (r < N)
          for (int b = r + 1; b < N; ++b) {
            double d, grx, gry, lapr, u, u1, u2;
            pair_dist(p, pc[2 * r] - pc[2 * b], pc[2 * r + 1] - pc[2 * b + 1], d, grx, gry, lapr);
            jastrow_u(p, sc[r] == sc[b], d, u, u1, u2);
            Ja += u;
          }
        J = group_sum_d<NR>(Ja);
      }
      if (out_lp && r == 0) {
        out_lp[2 * cfg] = L.re + J;
        out_lp[2 * cfg + 1] = L.im;
      }
    }
    if (off.dlw >= 0 && r == 0) {
      T* Ore = O + w * ldo;
      T* Oim = O + (W_total + w) * ldo;
      Ore[off.dlw + k] = (T)pk.re;
      Oim[off.dlw + k] = (T)pk.im;
      Ore[off.dlw + K + k] = (T)(-pk.im);
      Oim[off.dlw + K + k] = (T)pk.re;
    }
    // backflow cotangent of row i = q of M^-1: [Re, Im] channels of dL/dbf[k,a,i]
    T* y0 = ybar + (cfg * 2 + 0) * (int64_t)N * rowlen;
    T* y1 = ybar + (cfg * 2 + 1) * (int64_t)N * rowlen;
    if (f.zero) {
      // singular M: no inverse; the cotangent of this determinant is NaN, lane r < N taking row r
      if (r < N)
        for (int a = 0; a < N; ++a) {
          const int64_t col = (int64_t)r * rowlen + k * N + a;
          y0[col] = y0[col + KN] = y1[col] = y1[col + KN] = (T)NAN;
        }
    } else if (f.used) {
      const int i = f.q;
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int a = rowdet_pk(f, s);
        if (s >= N || a >= N) continue;
        const cd g = envc[a * EP + i] * (pk * f.m[s]);
        const int64_t col = (int64_t)i * rowlen + k * N + a;
        y0[col] = (T)g.re;
        y0[col + KN] = (T)(-g.im);
        y1[col] = (T)g.im;
        y1[col + KN] = (T)g.re;
      }
    }
  }
  if (p.use_jastrow) {
    const int nk = p.jastrow_kind == 0 ? p.n_knots : 1;
    for (int e = tid; e < ng * 2 * nk; e += nthr) {
      const int c = e / (2 * nk), q = e % (2 * nk);
      double v;
      const int64_t o = jastrow_o_entry(p, off, N, pos + c * 2 * NR, spin + c * NR, q, v);
      if (o < 0) continue;
      const int64_t w = w0 + cfg0 + c;
      O[w * ldo + o] = (T)v;
      O[(W_total + w) * ldo + o] = (T)0;
    }
  }
}

template <typename T>
int launch_det_bwd(int64_t n_cfg, const DetParams& p, const double* R, const int32_t* S, int64_t w0, const T* bf,
                   T* ybar, double* out_lp, T* O, int64_t ldo, int64_t W_total, const DetBwdOffsets& off,
                   hipStream_t st) {
  if (n_cfg == 0) return NQS_OK;
  const int N = p.N, NN = N * N;
  if (N <= 32 && p.K <= 16 && off.dW < 0) {
    const int NR = N <= 16 ? 16 : 32, GW = 64 / NR;
    auto bytes = [&](int G) { return rowdet_lds(NR, G, N, p.n_g, 0) + (size_t)G * 16 * sizeof(cd); };
    int G = std::max(1, std::min(4, 16 / p.K));
    while (G > 1 && bytes(G) > 64 * 1024) --G;
    const int waves = std::min(16 / GW, (G * p.K + GW - 1) / GW);
    const size_t lds = bytes(G);
    if (lds <= 64 * 1024) {
      const JetGeom j{n_cfg, N, 1, 0, 0, 0};
      const dim3 grid((unsigned)cdiv(n_cfg, G)), block(64 * waves);
      if (NR == 16)
        hipLaunchKernelGGL((k_det_bwd_rows<T, 16>), grid, block, lds, st, j, p, G, R, S, w0, bf, ybar, out_lp, O, ldo,
                           W_total, off);
      else
        hipLaunchKernelGGL((k_det_bwd_rows<T, 32>), grid, block, lds, st, j, p, G, R, S, w0, bf, ybar, out_lp, O, ldo,
                           W_total, off);
      NQS_LAUNCH_CHECK();
      return NQS_OK;
    }
  }
  size_t lds = sizeof(cd) * (size_t)(4 * NN + p.K) + sizeof(double) * 2 * N + sizeof(int) * (2 * N + 4);
  lds = round_up((int64_t)lds, 16);
  NQS_REQUIRE(lds <= 160 * 1024, NQS_ERR_UNSUPPORTED, "determinant backward exceeds LDS");
  hipLaunchKernelGGL(k_det_bwd<T>, dim3((unsigned)n_cfg), dim3(256), lds, st, p, R, S, w0, bf, ybar, out_lp, O, ldo,
                     W_total, off);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}
template int launch_det_bwd<double>(int64_t, const DetParams&, const double*, const int32_t*, int64_t, const double*,
                                    double*, double*, double*, int64_t, int64_t, const DetBwdOffsets&, hipStream_t);
template int launch_det_bwd<float>(int64_t, const DetParams&, const double*, const int32_t*, int64_t, const float*,
                                   float*, double*, float*, int64_t, int64_t, const DetBwdOffsets&, hipStream_t);

}  // namespace nqs
