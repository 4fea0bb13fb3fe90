// Standalone material-point update: the C-ABI behind MaterialModel.update
// (src/fem/materials/data.py:68-97) for J2ViscoplasticMaterial,
// LinearElasticMaterial and PhasedMaterialModel.
//
// One thread per point over SoA arrays ([component][point]) so that every
// load and store of a wavefront is one contiguous 512 B (FP64) run per row.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "common.hpp"

namespace cia {

static thread_local std::string g_last_error;

void set_error(const char* what, hipError_t err, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(err), file, line);
  g_last_error = buf;
}

void set_error_msg(const std::string& msg) { g_last_error = msg; }

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }

cia_status device_ctx(DeviceCtx** out) {
  static std::mutex mu;
  static std::vector<DeviceCtx*> ctxs;
  int dev = 0;
  CIA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)ctxs.size() <= dev) ctxs.resize(dev + 1, nullptr);
  if (!ctxs[dev]) {
    auto* c = new DeviceCtx();
    CIA_HIP(hipMalloc(&c->tab, sizeof(PhaseTable)));
    CIA_HIP(hipMalloc(&c->first_fail, sizeof(unsigned long long)));
    CIA_HIP(hipMalloc(&c->rec, sizeof(cia_point_failure)));
    ctxs[dev] = c;
  }
  *out = ctxs[dev];
  return CIA_OK;
}

template <class T, bool EXTRA>
__global__ void __launch_bounds__(kBlock)
    k_material_update(int64_t n, const PhaseTable* __restrict__ tab, Settings S, double dt,
                      const int32_t* __restrict__ phase_id, const T* __restrict__ strain,
                      const T* __restrict__ rate, const T* __restrict__ state_in, T* __restrict__ stress,
                      T* __restrict__ state_out, T* __restrict__ tangent, T* __restrict__ fe,
                      T* __restrict__ diss, T* __restrict__ diag, unsigned long long* first_fail) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ph = phase_id ? phase_id[p] : 0;
  const PhaseConst P = tab->p[ph];
  T eps[6], epsp[6], rt[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) eps[i] = strain[i * n + p];
  if (rate) {
#pragma unroll
    for (int i = 0; i < 6; ++i) rt[i] = rate[i * n + p];
  }
  T qbar = (T)0, gam = (T)0;
  const bool j2 = P.kind == CIA_MAT_J2_VISCOPLASTIC;
  if (j2) {
#pragma unroll
    for (int i = 0; i < 6; ++i) epsp[i] = state_in[i * n + p];
    qbar = state_in[6 * n + p];
    gam = state_in[7 * n + p];
  } else {
#pragma unroll
    for (int i = 0; i < 6; ++i) epsp[i] = (T)0;
  }
  PointOut<T> o;
  PointExtra<T> ex;
  const int st = material_point<T, EXTRA>(P, S, dt, eps, rate ? rt : nullptr, epsp, qbar, gam, o,
                                          EXTRA ? &ex : nullptr, nullptr);
  if (st != CIA_PT_OK) {
    atomicMin(first_fail, (unsigned long long)p);
    return;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) stress[i * n + p] = o.sig[i];
  if (j2) {
#pragma unroll
    for (int i = 0; i < 6; ++i) state_out[i * n + p] = o.epsp[i];
    state_out[6 * n + p] = o.qbar;
    state_out[7 * n + p] = o.gam;
  } else if (state_out != state_in && state_in != nullptr) {
#pragma unroll
    for (int i = 0; i < 8; ++i) state_out[i * n + p] = state_in[i * n + p];
  }
  if (tangent) {
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) tangent[(i * 6 + j) * n + p] = tangent_entry(o.tan, i, j);
  }
  if (EXTRA) {
    if (fe) fe[p] = ex.psi;
    if (diss) diss[p] = ex.diss;
    if (diag) {
#pragma unroll
      for (int i = 0; i < CIA_N_DIAG; ++i) diag[i * n + p] = ex.d[i];
    }
  }
}

// Re-run one point with the failure record attached (the first failing point).
template <class T>
__global__ void k_material_diagnose(int64_t n, int64_t p, const PhaseTable* __restrict__ tab,
                                    Settings S, double dt, const int32_t* __restrict__ phase_id,
                                    const T* __restrict__ strain, const T* __restrict__ rate,
                                    const T* __restrict__ state_in, cia_point_failure* rec) {
  const int ph = phase_id ? phase_id[p] : 0;
  const PhaseConst P = tab->p[ph];
  T eps[6], epsp[6], rt[6];
  for (int i = 0; i < 6; ++i) {
    eps[i] = strain[i * n + p];
    rt[i] = rate ? rate[i * n + p] : (T)0;
    epsp[i] = (P.kind == CIA_MAT_J2_VISCOPLASTIC) ? state_in[i * n + p] : (T)0;
  }
  T qbar = (P.kind == CIA_MAT_J2_VISCOPLASTIC) ? state_in[6 * n + p] : (T)0;
  T gam = (P.kind == CIA_MAT_J2_VISCOPLASTIC) ? state_in[7 * n + p] : (T)0;
  cia_point_failure r;
  memset(&r, 0, sizeof(r));
  PointOut<T> o;
  PointExtra<T> ex;
  const int st = material_point<T, true>(P, S, dt, eps, rate ? rt : nullptr, epsp, qbar, gam, o, &ex,
                                         &r);
  r.point_id = p;
  r.reason = st;
  *rec = r;
}

template <class T>
cia_status material_update(void* stream, int64_t n, const cia_material* phases, int32_t n_phases,
                           const int32_t* phase_id, const T* strain, const T* rate,
                           const T* state_in, double dt, const cia_update_settings* settings,
                           uint32_t flags, T* stress, T* state_out, T* tangent, T* fe, T* diss,
                           T* diag, cia_point_failure* failure) {
  CIA_ARG(n >= 0, "n must be non-negative");
  CIA_ARG(phases && n_phases >= 1 && n_phases <= CIA_MAX_PHASES, "invalid phase table");
  CIA_ARG(settings, "settings is NULL");
  CIA_ARG(strain && stress, "strain/stress are NULL");
  if (failure) {
    memset(failure, 0, sizeof(*failure));
    failure->point_id = -1;
  }
  bool any_j2 = false;
  for (int i = 0; i < n_phases; ++i) any_j2 |= phases[i].kind == CIA_MAT_J2_VISCOPLASTIC;
  CIA_ARG(!any_j2 || (state_in && state_out), "J2 phases need state_in/state_out");
  CIA_ARG(!(flags & CIA_OUT_TANGENT) || tangent, "CIA_OUT_TANGENT without tangent buffer");
  CIA_ARG(!(flags & CIA_OUT_ENERGY) || (fe && diss), "CIA_OUT_ENERGY without buffers");
  CIA_ARG(!(flags & CIA_OUT_DIAGNOSTICS) || diag, "CIA_OUT_DIAGNOSTICS without buffer");
  if (n == 0) return CIA_OK;
  DeviceCtx* ctx = nullptr;
  cia_status s = device_ctx(&ctx);
  if (s != CIA_OK) return s;
  hipStream_t st = (hipStream_t)stream;
  PhaseTable tab{};
  tab.n = n_phases;
  for (int i = 0; i < n_phases; ++i) tab.p[i] = make_phase(phases[i]);
  const Settings S = make_settings(*settings);
  CIA_HIP(hipMemcpyAsync(ctx->tab, &tab, sizeof(tab), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, ctx->first_fail, ~0ull);
  const bool extra = (flags & (CIA_OUT_ENERGY | CIA_OUT_DIAGNOSTICS)) != 0;
  T* tan_out = (flags & CIA_OUT_TANGENT) ? tangent : nullptr;
  if (extra)
    hipLaunchKernelGGL((k_material_update<T, true>), dim3(grid_for(n)), dim3(kBlock), 0, st, n,
                       ctx->tab, S, dt, phase_id, strain, rate, state_in, stress, state_out,
                       tan_out, (flags & CIA_OUT_ENERGY) ? fe : nullptr,
                       (flags & CIA_OUT_ENERGY) ? diss : nullptr,
                       (flags & CIA_OUT_DIAGNOSTICS) ? diag : nullptr, ctx->first_fail);
  else
    hipLaunchKernelGGL((k_material_update<T, false>), dim3(grid_for(n)), dim3(kBlock), 0, st, n,
                       ctx->tab, S, dt, phase_id, strain, rate, state_in, stress, state_out,
                       tan_out, nullptr, nullptr, nullptr, ctx->first_fail);
  CIA_HIP(hipGetLastError());
  unsigned long long ff = 0;
  CIA_HIP(hipMemcpyAsync(&ff, ctx->first_fail, sizeof(ff), hipMemcpyDeviceToHost, st));
  CIA_HIP(hipStreamSynchronize(st));
  if (ff != ~0ull) {
    // Diagnostics computed with EXTRA so that the increment-diagnostic
    // failure (non-finite previous stress) is reproduced as well.
    hipLaunchKernelGGL(k_material_diagnose<T>, dim3(1), dim3(1), 0, st, n, (int64_t)ff, ctx->tab, S,
                       dt, phase_id, strain, rate, state_in, ctx->rec);
    CIA_HIP(hipGetLastError());
    cia_point_failure rec;
    CIA_HIP(hipMemcpyAsync(&rec, ctx->rec, sizeof(rec), hipMemcpyDeviceToHost, st));
    CIA_HIP(hipStreamSynchronize(st));
    if (failure) *failure = rec;
    return CIA_ERR_POINT_FAILURE;
  }
  return CIA_OK;
}

}  // namespace cia

extern "C" {

int32_t cia_abi_version(void) { return 1; }

cia_status cia_last_error(char* buffer, int32_t capacity) {
  if (!buffer || capacity <= 0) return CIA_ERR_INVALID_ARGUMENT;
  snprintf(buffer, (size_t)capacity, "%s", cia::g_last_error.c_str());
  return CIA_OK;
}

cia_status cia_device_name(char* buffer, int32_t capacity) {
  if (!buffer || capacity <= 0) return CIA_ERR_INVALID_ARGUMENT;
  int dev = 0;
  CIA_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CIA_HIP(hipGetDeviceProperties(&prop, dev));
  snprintf(buffer, (size_t)capacity, "%s (%s)", prop.name, prop.gcnArchName);
  return CIA_OK;
}

cia_status cia_material_update_f64(void* stream, int64_t n, const cia_material* phases,
                                   int32_t n_phases, const int32_t* phase_id, const double* strain,
                                   const double* strain_rate, const double* state_in,
                                   double time_step, const cia_update_settings* settings,
                                   uint32_t flags, double* stress, double* state_out,
                                   double* tangent, double* free_energy, double* dissipation,
                                   double* diagnostics, cia_point_failure* failure) {
  return cia::material_update<double>(stream, n, phases, n_phases, phase_id, strain, strain_rate,
                                      state_in, time_step, settings, flags, stress, state_out,
                                      tangent, free_energy, dissipation, diagnostics, failure);
}

cia_status cia_material_update_f32(void* stream, int64_t n, const cia_material* phases,
                                   int32_t n_phases, const int32_t* phase_id, const float* strain,
                                   const float* strain_rate, const float* state_in,
                                   double time_step, const cia_update_settings* settings,
                                   uint32_t flags, float* stress, float* state_out, float* tangent,
                                   float* free_energy, float* dissipation, float* diagnostics,
                                   cia_point_failure* failure) {
  return cia::material_update<float>(stream, n, phases, n_phases, phase_id, strain, strain_rate,
                                     state_in, time_step, settings, flags, stress, state_out,
                                     tangent, free_energy, dissipation, diagnostics, failure);
}

}  // extern "C"
