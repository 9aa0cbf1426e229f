// pdr_ens.hip -- batched ensembles of small boxes: many independent
// realisations of uneven sizes advanced together.
//
// A 32^3 box gives a pass a few dozen workgroups, so stepping each member on
// its own leaves most of the 256 CUs idle and pays 15 kernel launches per step
// per member.  Here every pass of an RK3 stage is one launch over the tiles of
// all members (per-member work descriptors, pdr_kernels2.hip), and the 15
// launches of a step are captured once into a HIP graph and replayed, so a
// step costs one graph launch for the whole ensemble.  Each member runs the
// device bodies that pdr_step runs for it, with the same arguments; they are
// compiled into other kernels, where the compiler may contract multiply-adds
// differently, so the results agree with pdr_step to rounding.
//
// Scope: members of one precision with the specialised kernels, cubic, line
// lengths 32..256, isotropic with RK3_TIMESTEP = FIXED, no forcing and no
// tracers (the decaying-ensemble case).  Those keep every kernel argument of
// the step constant, which is what makes the step a replayable graph; shear
// frames (per-stage coefficients), CFL steps (a host read of the maximum),
// forcing (per-step draws) and tracers are rejected with PDR_ERR_UNSUPPORTED.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "pdr_internal.hpp"

struct pdr_ensemble {
  std::vector<pdr_ctx*> m;
  int prec = 0;
  hipStream_t stream = nullptr;
  pdr::EnsMember* d_mem = nullptr;
  unsigned nblk[5] = {0, 0, 0, 0, 0};
  size_t lds[5] = {0, 0, 0, 0, 0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int fpb_y = 0, fpb_yf = 0;  // 0: each member's own choice
};

namespace {

using pdr::set_error;

const double kBeta[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
const double kGamma[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

void ens_release(pdr_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

int ens_build(pdr_ensemble* e) {
  ens_release(e);
  const int nm = int(e->m.size());
  std::vector<pdr::EnsMember> h(static_cast<size_t>(nm));
  for (int q = 0; q < 5; ++q) {
    e->nblk[q] = 0;
    e->lds[q] = 0;
  }
  for (int k = 0; k < nm; ++k) {
    pdr_ctx* c = e->m[size_t(k)];
    pdr::EnsMember& d = h[size_t(k)];
    std::memset(&d, 0, sizeof(d));
    d.g = c->g;
    d.cf = pdr::make_coef(c);
    d.n = c->g.nx;
    d.fpb_y = e->fpb_y ? e->fpb_y : c->fpb_y;
    d.fpb_yf = e->fpb_yf ? e->fpb_yf : c->fpb_yf;
    unsigned blocks[5];
    if (pdr::ens_tiles(c->prec, d.n, c->g, d.fpb_y, d.fpb_yf, d.nt, blocks) != 0)
      return set_error(PDR_ERR_UNSUPPORTED, "ensemble: unsupported line length");
    for (int q = 0; q < 5; ++q) {
      d.start[q] = e->nblk[q];
      e->nblk[q] += blocks[q];
      e->lds[q] = std::max(e->lds[q], pdr::ens_lds(c->prec, d.n, q));
    }
    d.dt = c->p.rk3_dt;
    d.norm = 1.0 / (double(c->g.nx) * double(c->g.ny) * double(c->g.nz));
    for (int q = 0; q < 3; ++q) {
      d.u[q] = c->u[q];
      d.rk[q] = c->rk[q];
    }
    d.w1 = c->w1;
    d.w2 = c->w2;
    d.umax = c->umax;
    d.twz = c->tab.twz;
    d.twy = c->tab.twy;
    d.twx = c->tab.twx;
  }
  if (!e->d_mem && hipMalloc(&e->d_mem, sizeof(pdr::EnsMember) * size_t(nm)) != hipSuccess) {
    (void)hipGetLastError();
    e->d_mem = nullptr;
    return set_error(PDR_ERR_NOMEM, "ensemble: device allocation failed");
  }
  size_t lmax = 0;
  for (int q = 0; q < 5; ++q) lmax = std::max(lmax, e->lds[q]);
  hipError_t r = pdr::ens_prepare(e->prec, lmax);
  if (r == hipSuccess)
    r = hipMemcpy(e->d_mem, h.data(), sizeof(pdr::EnsMember) * size_t(nm), hipMemcpyHostToDevice);
  // capture one RK3 step: 3 stages x 5 batched passes
  if (r == hipSuccess) r = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
  if (r == hipSuccess) {
    hipError_t lr = hipSuccess;
    for (int st = 0; st < 3 && lr == hipSuccess; ++st)
      for (int q = 0; q < 5 && lr == hipSuccess; ++q)
        lr = pdr::launch_ens_pass(e->prec, q, e->d_mem, nm, e->nblk[q], e->lds[q], kBeta[st], kGamma[st],
                                  e->stream);
    r = hipStreamEndCapture(e->stream, &e->graph);
    if (lr != hipSuccess) r = lr;
  }
  if (r == hipSuccess) r = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
  if (r != hipSuccess) {
    ens_release(e);
    return set_error(PDR_ERR_HIP, std::string("ensemble set-up: ") + hipGetErrorString(r));
  }
  return PDR_OK;
}

}  // namespace

extern "C" {

int pdr_ensemble_create(pdr_ensemble** out, pdr_ctx* const* members, int32_t count) {
  if (!out || !members || count < 1) return set_error(PDR_ERR_ARG, "pdr_ensemble_create: bad argument");
  *out = nullptr;
  const int prec = members[0] ? members[0]->prec : 0;
  for (int32_t k = 0; k < count; ++k) {
    const pdr_ctx* c = members[k];
    if (!c) return set_error(PDR_ERR_ARG, "pdr_ensemble_create: null member");
    for (int32_t q = 0; q < k; ++q)
      if (members[q] == c) return set_error(PDR_ERR_ARG, "pdr_ensemble_create: member listed twice");
    const bool ok = c->prec == prec && c->spec && c->fuse_zf && c->g.zc == c->g.nz && c->g.nx == c->g.ny &&
                    c->g.ny == c->g.nz && pdr::ens_size_supported(c->g.nx) &&
                    c->p.f_type == PDR_F_ISOTROPIC && c->p.rk3_timestep == PDR_RK3_TIMESTEP_FIXED &&
                    !c->forcing && c->np == 0 && c->p.dealiasing == PDR_DEALIASING_TWOTHIRD;
    if (!ok)
      return set_error(PDR_ERR_UNSUPPORTED,
                       "pdr_ensemble_create: members must share the precision and be cubic (32..256 points), "
                       "isotropic, FIXED time step, 2/3 dealiasing, unforced and without tracers");
  }
  pdr_ensemble* e = new pdr_ensemble();
  e->m.assign(members, members + count);
  e->prec = prec;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete e;
    return set_error(PDR_ERR_HIP, "pdr_ensemble_create: stream creation failed");
  }
  int rc = ens_build(e);
  if (rc != PDR_OK) {
    pdr_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return PDR_OK;
}

int pdr_ensemble_set_split(pdr_ensemble* e, int32_t fpb_y, int32_t fpb_yf) {
  if (!e) return set_error(PDR_ERR_ARG, "null ensemble");
  const bool ok = (fpb_y == 0 || fpb_y == 1 || fpb_y == 2 || fpb_y == 3 || fpb_y == 6) &&
                  (fpb_yf == 0 || fpb_yf == 1 || fpb_yf == 3);
  if (!ok) return set_error(PDR_ERR_ARG, "pdr_ensemble_set_split: fields per workgroup must divide 6 / 3");
  (void)hipStreamSynchronize(e->stream);
  e->fpb_y = fpb_y;
  e->fpb_yf = fpb_yf;
  return ens_build(e);
}

// Times `steps` graph steps for each split of the y passes over workgroups
// (fields per workgroup 1, 2, 3, 6 for the inverse pass and 1, 3 for the
// forward pass), keeps the fastest, and restores every member's state and RK
// registers, so the ensemble is where it was before the call.
int pdr_ensemble_autotune(pdr_ensemble* e, int32_t steps, int32_t* best_y, int32_t* best_yf) {
  if (!e || steps < 1) return set_error(PDR_ERR_ARG, "pdr_ensemble_autotune: bad argument");
  for (pdr_ctx* c : e->m)
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return set_error(PDR_ERR_HIP, "pdr_ensemble_autotune: member stream failed");
  // save the state
  size_t total = 0;
  for (pdr_ctx* c : e->m) total += 6 * size_t(c->g.mw) * c->csz;
  char* save = nullptr;
  if (hipMalloc(&save, total) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(PDR_ERR_NOMEM, "pdr_ensemble_autotune: device allocation failed");
  }
  auto copy_state = [&](bool to_save) {
    size_t off = 0;
    hipError_t r = hipSuccess;
    for (pdr_ctx* c : e->m) {
      const size_t b = size_t(c->g.mw) * c->csz;
      void* f[6] = {c->u[0], c->u[1], c->u[2], c->rk[0], c->rk[1], c->rk[2]};
      for (void* a : f) {
        if (r == hipSuccess)
          r = to_save ? hipMemcpyAsync(save + off, a, b, hipMemcpyDeviceToDevice, e->stream)
                      : hipMemcpyAsync(a, save + off, b, hipMemcpyDeviceToDevice, e->stream);
        off += b;
      }
    }
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    return r;
  };
  hipError_t r = copy_state(true);
  hipEvent_t a = nullptr, b = nullptr;
  if (r == hipSuccess) r = hipEventCreate(&a);
  if (r == hipSuccess) r = hipEventCreate(&b);
  const int ys[4] = {1, 2, 3, 6}, yfs[2] = {1, 3};
  float best = 1e30f;
  int by = e->fpb_y, byf = e->fpb_yf;
  int rc = PDR_OK;
  for (int iy = 0; iy < 4 && r == hipSuccess && rc == PDR_OK; ++iy)
    for (int iyf = 0; iyf < 2 && r == hipSuccess && rc == PDR_OK; ++iyf) {
      e->fpb_y = ys[iy];
      e->fpb_yf = yfs[iyf];
      rc = ens_build(e);
      if (rc != PDR_OK) break;
      r = hipGraphLaunch(e->exec, e->stream);  // warm-up
      if (r == hipSuccess) r = hipEventRecord(a, e->stream);
      for (int32_t s = 0; s < steps && r == hipSuccess; ++s) r = hipGraphLaunch(e->exec, e->stream);
      if (r == hipSuccess) r = hipEventRecord(b, e->stream);
      if (r == hipSuccess) r = hipEventSynchronize(b);
      float ms = 0.0f;
      if (r == hipSuccess) r = hipEventElapsedTime(&ms, a, b);
      if (r == hipSuccess && ms < best) {
        best = ms;
        by = ys[iy];
        byf = yfs[iyf];
      }
    }
  if (r == hipSuccess) r = copy_state(false);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipFree(save);
  if (rc != PDR_OK) return rc;
  if (r != hipSuccess) return set_error(PDR_ERR_HIP, std::string("pdr_ensemble_autotune: ") + hipGetErrorString(r));
  e->fpb_y = by;
  e->fpb_yf = byf;
  if (best_y) *best_y = by;
  if (best_yf) *best_yf = byf;
  return ens_build(e);
}

int pdr_ensemble_run(pdr_ensemble* e, int32_t tstep0, int32_t nsteps, double* simtime) {
  (void)tstep0;
  if (!e || nsteps < 0 || !simtime) return set_error(PDR_ERR_ARG, "pdr_ensemble_run: bad argument");
  for (pdr_ctx* c : e->m)  // work queued on the members' own streams comes first
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return set_error(PDR_ERR_HIP, "pdr_ensemble_run: member stream failed");
  for (int32_t s = 0; s < nsteps; ++s) {
    hipError_t r = hipGraphLaunch(e->exec, e->stream);
    if (r != hipSuccess) return set_error(PDR_ERR_HIP, std::string("pdr_ensemble_run: ") + hipGetErrorString(r));
    // g_RK3Calc: simtime = t + dt after the third stage (src/pg_rk3.f90:82-129)
    for (size_t k = 0; k < e->m.size(); ++k) {
      pdr_ctx* c = e->m[k];
      c->dt = c->p.rk3_dt;
      simtime[k] = simtime[k] + c->dt;
    }
  }
  if (hipStreamSynchronize(e->stream) != hipSuccess) return set_error(PDR_ERR_HIP, "pdr_ensemble_run: step failed");
  return PDR_OK;
}

int pdr_ensemble_destroy(pdr_ensemble* e) {
  if (!e) return PDR_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_release(e);
  if (e->d_mem) (void)hipFree(e->d_mem);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return PDR_OK;
}

}  // extern "C"
