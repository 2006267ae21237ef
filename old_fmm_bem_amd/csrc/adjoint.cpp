// adjoint.cpp -- the matvec's adjoint on the device: the transposed matvec
// (targets and sources exchanged), the gradients of L = w . (K q) with
// respect to the strengths and to both point sets, and the energy
// E = 1/2 q . K q with its gradients.
//
// For v_i = sum_j K(x_i, x_j) q_j (fmm/fmm.cxx:29, kernel/G_p2p.cxx:7):
//   (K^T w)_j = sum_i K(x_i, x_j) w_i: the same kernel with the roles of the
//     point sets exchanged -- G is symmetric; the target-normal kernel Gni
//     (kernel/Gni_p2p.cxx:7-32, (x_i - x_j).n_i) transposes into minus the
//     source-normal kernel Gnj (kernel/Gnj_p2p.cxx:7-32) carrying the target
//     normals, and Gnj into minus Gni carrying the source normals;
//   dL/dx_i = w_i grad_x phi(x_i), the field gradient L2P and P2P already
//     return (kernel/G_l2p.cxx:68-70);
//   dL/dx_j = q_j grad_y psi(x_j) with psi = K^T w (G depends on x_i - x_j
//     and is even, so the gradient in the source position is the gradient of
//     the transposed field at the source).
// The transposed problem runs in a second engine on the same root cube and
// depth whose targets are this handle's sources (key order) and whose
// sources are its targets; it is built at the first transposed evaluation
// after set_points.  Expansions are recomputed per evaluation, never stored.
#include <algorithm>
#include <cmath>

#include "engine.h"

namespace fmmx {

namespace {

constexpr int TPB = 256;
inline int nblk(long n) { return (int)((n + TPB - 1) / TPB); }

// dst[k] = src[perm[k]] for records of c floats (c = 2: complex pairs)
__global__ void k_gather_rec(int n, int c, const uint32_t* __restrict__ perm, const float* __restrict__ src,
                             float* __restrict__ dst) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t j = perm[k];
  for (int t = 0; t < c; t++) dst[(size_t)k * c + t] = src[(size_t)j * c + t];
}

// dst[perm[k]] = s * src[k] for records of c doubles
__global__ void k_scatter_rec(int n, int c, const uint32_t* __restrict__ perm, const double* __restrict__ src,
                              double s, double* __restrict__ dst) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t j = perm[k];
  for (int t = 0; t < c; t++) dst[(size_t)j * c + t] = s * src[(size_t)k * c + t];
}

// dx[a n + i] = s[i] * g_a[i], three components as consecutive blocks
__global__ void k_weight3(int n, const float* __restrict__ s, const double* __restrict__ g0,
                          const double* __restrict__ g1, const double* __restrict__ g2, double* __restrict__ dx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double w = (double)s[i];
  dx[i] = w * g0[i];
  dx[(size_t)n + i] = w * g1[i];
  dx[2 * (size_t)n + i] = w * g2[i];
}

// 1/2 sum_i q_i phi_i in two deterministic stages: per-block sums in a fixed
// tree order, then the block sums in order by one block
constexpr int RB = 256;
__global__ __launch_bounds__(RB) void k_dot_blocks(int n, const float* __restrict__ q, const double* __restrict__ v,
                                                   double* __restrict__ part) {
  __shared__ double s[RB];
  double a = 0;
  for (long i = (long)blockIdx.x * RB + threadIdx.x; i < n; i += (long)gridDim.x * RB) a += (double)q[i] * v[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ __launch_bounds__(RB) void k_sum_half(int nb, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double s[RB];
  double a = 0;
  for (int i = threadIdx.x; i < nb; i += RB) a += part[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = 0.5 * s[0];
}

const char* kernel_name(int neq) {
  return neq == FMMX_G ? "G" : neq == FMMX_GNI ? "Gni" : neq == FMMX_GNJ ? "Gnj" : neq == FMMX_YUKAWA ? "Yukawa"
                                                                                                       : "Helmholtz";
}

}  // namespace

int Engine::transposed(Engine*& t) {
  int rc;
  if (!tr) {
    tr = std::make_unique<Engine>();
    fmmx_config c = cfg;
    c.lmax = lmax;
    c.xmin = xmin;
    c.ymin = ymin;
    c.zmin = zmin;
    c.rd = rd;
    if ((rc = tr->init(&c, device, stream))) {
      tr.reset();
      return rc;
    }
    tr_valid = false;
  }
  Engine& T = *tr;
  T.lam = lam;
  T.hk = hk;
  for (int l = 0; l <= MAXL; l++) T.tune[l] = tune[l];
  if (!tr_valid) {
    if ((rc = T.set_domain(lmax, xmin, ymin, zmin, rd)) || (rc = T.set_groups(group_level))) return rc;
    // targets: this handle's sources; sources: its targets (key-ordered
    // device copies; the same arrays when both sides share one tree)
    auto nrm = [](const Side& s, const DevBuf& b) { return s.has_normals ? b.as<float>() : nullptr; };
    if ((rc = T.set_points(sj.n, sj.x.as<float>(), sj.y.as<float>(), sj.z.as<float>(), nrm(sj, sj.nx),
                           nrm(sj, sj.ny), nrm(sj, sj.nz), si.n, si.x.as<float>(), si.y.as<float>(),
                           si.z.as<float>(), nrm(si, si.nx), nrm(si, si.ny), nrm(si, si.nz), FMMX_DEVICE_PTRS)))
      return rc;
    tr_valid = true;
  }
  t = tr.get();
  return FMMX_OK;
}

int Engine::evaluate_transpose(int neq, const float* wi, double* uj, double* gxj, double* gyj, double* gzj,
                               int flags) {
  if (!have_points) return fail(FMMX_ESTATE, "transposed evaluation before set_points");
  if (neq != FMMX_G && neq != FMMX_GNI && neq != FMMX_GNJ && neq != FMMX_YUKAWA && neq != FMMX_HELMHOLTZ)
    return fail(FMMX_EINVAL, "neq must be 0, 1, 2, 10 or 20");
  const bool grad = gxj || gyj || gzj;
  if (grad && !(neq == FMMX_G || neq == FMMX_YUKAWA))
    return fail(FMMX_EUNSUPPORTED, std::string("no gradient for the transposed ") + kernel_name(neq) + " kernel");
  if (grad && !(gxj && gyj && gzj)) return fail(FMMX_EINVAL, "gradient outputs: all three or none");
  if (neq == FMMX_GNI && si.n > 0 && !si.has_normals) return fail(FMMX_EINVAL, "Gni needs target normals");
  if (neq == FMMX_GNJ && sj.n > 0 && !sj.has_normals) return fail(FMMX_EINVAL, "Gnj needs source normals");
  if ((si.n > 0 && !wi) || (sj.n > 0 && !uj)) return fail(FMMX_EINVAL, "w or u pointer is NULL");
  FMMX_HIP(hipSetDevice(device));
  Engine* T = nullptr;
  int rc = transposed(T);
  if (rc) return rc;
  const bool dev = flags & FMMX_DEVICE_PTRS;
  const int ni = si.n, nj = sj.n, c = neq == FMMX_HELMHOLTZ ? 2 : 1;
  const size_t nj1 = (size_t)std::max(nj, 1);
  // strengths w in this handle's target (key) order = the transposed
  // problem's caller order of its sources
  if ((rc = adj_w.ensure((size_t)std::max(ni, 1) * c * 4)) || (rc = adj_wk.ensure((size_t)std::max(ni, 1) * c * 4)))
    return rc;
  const float* w = wi;
  if (!dev && ni > 0) {
    FMMX_HIP(hipMemcpyAsync(adj_w.p, wi, (size_t)ni * c * 4, hipMemcpyHostToDevice, stream));
    w = adj_w.as<float>();
  }
  if (ni > 0) k_gather_rec<<<nblk(ni), TPB, 0, stream>>>(ni, c, si.perm.as<uint32_t>(), w, adj_wk.as<float>());
  for (int k = 0; k < 4; k++)
    if ((rc = adj_t[k].ensure(nj1 * c * 8))) return rc;
  const int neqT = neq == FMMX_GNI ? FMMX_GNJ : neq == FMMX_GNJ ? FMMX_GNI : neq;
  const double sign = (neq == FMMX_GNI || neq == FMMX_GNJ) ? -1.0 : 1.0;
  double* tk[4] = {adj_t[0].as<double>(), grad ? adj_t[1].as<double>() : nullptr,
                   grad ? adj_t[2].as<double>() : nullptr, grad ? adj_t[3].as<double>() : nullptr};
  if ((rc = T->evaluate(neqT, adj_wk.as<float>(), tk[0], tk[1], tk[2], tk[3], FMMX_DEVICE_PTRS))) return rc;
  // the transposed problem's outputs are in this handle's source key order
  double* outs[4] = {uj, gxj, gyj, gzj};
  for (int k = 0; k < 4 && nj > 0; k++) {
    if (!tk[k]) continue;
    double* dst = outs[k];
    if (!dev) {
      if ((rc = adj_t[4 + k].ensure(nj1 * c * 8))) return rc;
      dst = adj_t[4 + k].as<double>();
    }
    k_scatter_rec<<<nblk(nj), TPB, 0, stream>>>(nj, k == 0 ? c : 1, sj.perm.as<uint32_t>(), tk[k], sign, dst);
    if (!dev) FMMX_HIP(hipMemcpyAsync(outs[k], dst, (size_t)nj * (k == 0 ? c : 1) * 8, hipMemcpyDeviceToHost, stream));
  }
  FMMX_HIP(hipGetLastError());
  if (!dev) FMMX_HIP(hipStreamSynchronize(stream));
  return FMMX_OK;
}

int Engine::gradients(int neq, const float* qj, const float* wi, double* dq, double* dxi, double* dxj, int flags) {
  if (neq != FMMX_G && neq != FMMX_YUKAWA)
    return fail(FMMX_EUNSUPPORTED, "position gradients are provided for G (and Yukawa) only");
  if (!have_points) return fail(FMMX_ESTATE, "gradients before set_points");
  const int ni = si.n, nj = sj.n;
  if ((nj > 0 && !qj) || (ni > 0 && !wi)) return fail(FMMX_EINVAL, "q or w pointer is NULL");
  FMMX_HIP(hipSetDevice(device));
  const bool dev = flags & FMMX_DEVICE_PTRS;
  const size_t ni1 = (size_t)std::max(ni, 1), nj1 = (size_t)std::max(nj, 1);
  int rc;
  // device copies of the strengths the weights come from
  const float *q = qj, *w = wi;
  if (!dev) {
    if ((rc = adj_q.ensure(nj1 * 4)) || (rc = adj_w2.ensure(ni1 * 4))) return rc;
    if (nj > 0) FMMX_HIP(hipMemcpyAsync(adj_q.p, qj, (size_t)nj * 4, hipMemcpyHostToDevice, stream));
    if (ni > 0) FMMX_HIP(hipMemcpyAsync(adj_w2.p, wi, (size_t)ni * 4, hipMemcpyHostToDevice, stream));
    q = adj_q.as<float>();
    w = adj_w2.as<float>();
  }
  if (dxi && ni > 0) {
    // forward field gradient at the targets: dL/dx_i = w_i grad phi(x_i)
    for (int k = 8; k < 12; k++)
      if ((rc = adj_t[k].ensure(ni1 * 8))) return rc;
    if ((rc = evaluate(neq, q, adj_t[8].as<double>(), adj_t[9].as<double>(), adj_t[10].as<double>(),
                       adj_t[11].as<double>(), FMMX_DEVICE_PTRS)))
      return rc;
    double* dst = dxi;
    if (!dev) {
      if ((rc = adj_t[12].ensure(3 * ni1 * 8))) return rc;
      dst = adj_t[12].as<double>();
    }
    k_weight3<<<nblk(ni), TPB, 0, stream>>>(ni, w, adj_t[9].as<double>(), adj_t[10].as<double>(),
                                            adj_t[11].as<double>(), dst);
    if (!dev) FMMX_HIP(hipMemcpyAsync(dxi, dst, (size_t)ni * 24, hipMemcpyDeviceToHost, stream));
  }
  if ((dq || dxj) && nj > 0) {
    // transposed field psi = K^T w and its gradient at the sources:
    // dL/dq_j = psi_j, dL/dx_j = q_j grad psi(x_j)
    for (int k = 13; k < 17; k++)
      if ((rc = adj_t[k].ensure(nj1 * 8))) return rc;
    if ((rc = evaluate_transpose(neq, w, adj_t[13].as<double>(), adj_t[14].as<double>(), adj_t[15].as<double>(),
                                 adj_t[16].as<double>(), FMMX_DEVICE_PTRS)))
      return rc;
    if (dq) {
      const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
      FMMX_HIP(hipMemcpyAsync(dq, adj_t[13].p, (size_t)nj * 8, kind, stream));
    }
    if (dxj) {
      double* dst = dxj;
      if (!dev) {
        if ((rc = adj_t[17].ensure(3 * nj1 * 8))) return rc;
        dst = adj_t[17].as<double>();
      }
      k_weight3<<<nblk(nj), TPB, 0, stream>>>(nj, q, adj_t[14].as<double>(), adj_t[15].as<double>(),
                                              adj_t[16].as<double>(), dst);
      if (!dev) FMMX_HIP(hipMemcpyAsync(dxj, dst, (size_t)nj * 24, hipMemcpyDeviceToHost, stream));
    }
  }
  FMMX_HIP(hipGetLastError());
  if (!dev) FMMX_HIP(hipStreamSynchronize(stream));
  return FMMX_OK;
}

int Engine::energy(int neq, const float* qj, double* E, double* dEdq, double* dEdx, int flags) {
  if (neq != FMMX_G && neq != FMMX_YUKAWA) return fail(FMMX_EUNSUPPORTED, "the energy is provided for G (and Yukawa)");
  if (!have_points) return fail(FMMX_ESTATE, "energy before set_points");
  if (si.n != sj.n || si.x.p != sj.x.p)
    return fail(FMMX_EINVAL, "the energy needs targets = sources (one point set in set_points)");
  const int n = si.n;
  if (n > 0 && !qj) return fail(FMMX_EINVAL, "q pointer is NULL");
  FMMX_HIP(hipSetDevice(device));
  const bool dev = flags & FMMX_DEVICE_PTRS;
  const size_t n1 = (size_t)std::max(n, 1);
  int rc;
  const float* q = qj;
  if (!dev) {
    if ((rc = adj_q.ensure(n1 * 4))) return rc;
    if (n > 0) FMMX_HIP(hipMemcpyAsync(adj_q.p, qj, (size_t)n * 4, hipMemcpyHostToDevice, stream));
    q = adj_q.as<float>();
  }
  for (int k = 8; k < 12; k++)
    if ((rc = adj_t[k].ensure(n1 * 8))) return rc;
  // phi = K q with its gradient: dE/dq = phi, dE/dx_k = q_k grad phi(x_k)
  if ((rc = evaluate(neq, q, adj_t[8].as<double>(), adj_t[9].as<double>(), adj_t[10].as<double>(),
                     adj_t[11].as<double>(), FMMX_DEVICE_PTRS)))
    return rc;
  const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (dEdq && n > 0) FMMX_HIP(hipMemcpyAsync(dEdq, adj_t[8].p, (size_t)n * 8, kind, stream));
  if (dEdx && n > 0) {
    double* dst = dEdx;
    if (!dev) {
      if ((rc = adj_t[12].ensure(3 * n1 * 8))) return rc;
      dst = adj_t[12].as<double>();
    }
    k_weight3<<<nblk(n), TPB, 0, stream>>>(n, q, adj_t[9].as<double>(), adj_t[10].as<double>(),
                                           adj_t[11].as<double>(), dst);
    if (!dev) FMMX_HIP(hipMemcpyAsync(dEdx, dst, (size_t)n * 24, hipMemcpyDeviceToHost, stream));
  }
  if (E) {
    const int nb = std::max(1, std::min(1024, (n + RB - 1) / RB));
    if ((rc = adj_t[18].ensure((size_t)(nb + 1) * 8))) return rc;
    double* part = adj_t[18].as<double>();
    k_dot_blocks<<<nb, RB, 0, stream>>>(n, q, adj_t[8].as<double>(), part);
    k_sum_half<<<1, RB, 0, stream>>>(nb, part, part + nb);
    if (dev) {
      FMMX_HIP(hipMemcpyAsync(E, part + nb, 8, hipMemcpyDeviceToDevice, stream));
    } else {
      FMMX_HIP(hipMemcpyAsync(E, part + nb, 8, hipMemcpyDeviceToHost, stream));
    }
  }
  FMMX_HIP(hipGetLastError());
  if (!dev) FMMX_HIP(hipStreamSynchronize(stream));
  return FMMX_OK;
}

}  // namespace fmmx
