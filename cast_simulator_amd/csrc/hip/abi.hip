// C ABI entry points of libcast_hip.so (declared in include/cast_hip.h).
//
// Every entry point is an exception firewall: implementation errors are
// C++ exceptions that are caught here and turned into (return 1 / null,
// err_buf message), as src/cpp/cuda/ffi_cuda.cpp:138-144 does.

#include "cast_hip.h"
#include "common.hpp"

#include <map>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

namespace cast_hip {

void ensure_device() {
  int n = 0;
  const hipError_t rc = hipGetDeviceCount(&n);
  if (rc != hipSuccess || n <= 0)
    throw Error(std::string("no HIP device available (hipGetDeviceCount: ") +
                hipGetErrorString(rc) + ")");
}

const DeviceProps &device_props() {
  static std::once_flag once;
  static DeviceProps props;
  std::call_once(once, [] {
    ensure_device();
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t p{};
    check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
    props.device = dev;
    props.n_cu = p.multiProcessorCount;
    props.warp = p.warpSize;
    props.lds_per_block = p.sharedMemPerBlock;
    props.l2_bytes = p.l2CacheSize;
    props.arch = p.gcnArchName;
    props.name = p.name;
  });
  return props;
}

namespace {

template <typename F> int guard(char *err, size_t len, F &&f) {
  try {
    f();
    clear_err(err, len);
    return 0;
  } catch (const std::exception &e) {
    write_err(err, len, e.what());
  } catch (...) {
    write_err(err, len, "unknown C++ exception");
  }
  return 1;
}

// Every statevector kernel moves 16-byte vectors (one f64 amplitude, two f32
// ones); allocations from cast_hip_sv_alloc are 256-byte aligned.
void check_aligned(uint64_t dptr) {
  if (dptr % 16 != 0)
    throw Error("statevector pointer must be 16-byte aligned");
}

void check_sv_args(uint64_t dptr, size_t n_elements, uint8_t precision) {
  check_precision(precision);
  if (dptr == 0)
    throw Error("null device pointer");
  if (n_elements == 0)
    throw Error("n_elements must be positive");
  check_aligned(dptr);
}

// Matrix scalar classification of src/cpp/cuda/cuda_gen.cpp:67-112: the
// tolerances are divided by K, zeros are tested first, then +1, then -1.
double snap(double v, double ztol, double otol, double K, bool fp32) {
  if (ztol > 0.0 && std::fabs(v) < ztol / K)
    return 0.0;
  if (otol > 0.0 && std::fabs(v - 1.0) < otol / K)
    return 1.0;
  if (otol > 0.0 && std::fabs(v + 1.0) < otol / K)
    return -1.0;
  return fp32 ? double(float(v)) : v;
}

GatePlan *build_plan(const cast_hip_kernel_gen_spec_t *spec, const cast_complex64_t *matrix,
                     size_t matrix_len, const uint32_t *qubits, size_t n_qubits) {
  if (spec == nullptr)
    throw Error("spec must not be null");
  if (matrix == nullptr)
    throw Error("matrix pointer must not be null");
  if (qubits == nullptr || n_qubits == 0)
    throw Error("qubits must not be null/empty");
  if (n_qubits > 6)
    throw Error("gate size k=" + std::to_string(n_qubits) + " exceeds the maximum of 6");
  const size_t K = size_t(1) << n_qubits;
  if (matrix_len != K * K)
    throw Error("matrix_len must equal (2^n_qubits)^2");
  if (spec->precision != CAST_PRECISION_F32 && spec->precision != CAST_PRECISION_F64)
    throw Error("spec.precision must be F32 or F64");
  for (size_t b = 0; b < n_qubits; ++b) {
    if (qubits[b] > 40)
      throw Error("qubit index " + std::to_string(qubits[b]) + " out of range");
    if (b > 0 && qubits[b] <= qubits[b - 1])
      throw Error("qubits must be strictly ascending");
  }
  if (spec->variant > CAST_HIP_VARIANT_MFMA)
    throw Error("unknown variant " + std::to_string(spec->variant));

  auto plan = std::make_unique<GatePlan>();
  plan->k = uint32_t(n_qubits);
  plan->precision = uint8_t(spec->precision);
  for (size_t b = 0; b < n_qubits; ++b)
    plan->q[b] = qubits[b];
  const bool fp32 = spec->precision == CAST_PRECISION_F32;

  std::vector<double> m(2 * K * K);
  uint64_t nnz = 0;
  for (size_t i = 0; i < K * K; ++i) {
    m[2 * i] = snap(matrix[i].re, spec->ztol, spec->otol, double(K), fp32);
    m[2 * i + 1] = snap(matrix[i].im, spec->ztol, spec->otol, double(K), fp32);
    nnz += (m[2 * i] != 0.0) + (m[2 * i + 1] != 0.0);
  }
  plan->scalar_nnz = nnz;

  // Structure: monomial when every row has at most one nonzero entry.
  bool monomial = true;
  uint32_t n_active = 0;
  for (size_t r = 0; r < K && monomial; ++r) {
    int col = -1;
    for (size_t c = 0; c < K; ++c) {
      if (m[2 * (r * K + c)] != 0.0 || m[2 * (r * K + c) + 1] != 0.0) {
        if (col >= 0) {
          monomial = false;
          break;
        }
        col = int(c);
      }
    }
    if (!monomial)
      break;
    double re = 0.0, im = 0.0;
    if (col >= 0) {
      re = m[2 * (r * K + col)];
      im = m[2 * (r * K + col) + 1];
    } else {
      col = int(r); // all-zero row: output 0
    }
    const bool identity_row = size_t(col) == r && re == 1.0 && im == 0.0;
    if (!identity_row) {
      plan->act_row[n_active] = uint8_t(r);
      plan->act_col[n_active] = uint8_t(col);
      plan->act_re[n_active] = re;
      plan->act_im[n_active] = im;
      ++n_active;
    }
  }
  plan->n_active = n_active;

  uint32_t variant = spec->variant;
  if (variant == CAST_HIP_VARIANT_AUTO)
    variant = monomial ? CAST_HIP_VARIANT_MONOMIAL : CAST_HIP_VARIANT_DENSE;
  if (variant == CAST_HIP_VARIANT_MONOMIAL && !monomial)
    throw Error("variant MONOMIAL requested for a matrix with >1 nonzero in a row");
  if (variant == CAST_HIP_VARIANT_TILE && n_qubits > 5)
    variant = CAST_HIP_VARIANT_DENSE; // wave tiles exist for k <= 5
  if (variant == CAST_HIP_VARIANT_MFMA && n_qubits < 4)
    variant = CAST_HIP_VARIANT_DENSE; // matrix-core tiles exist for k = 4, 5, 6
  plan->variant = variant;
  plan->forced = spec->variant != CAST_HIP_VARIANT_AUTO;
  if ((spec->flags & CAST_HIP_FLAG_MFMA_SPLIT) && (spec->flags & CAST_HIP_FLAG_MFMA_F32))
    throw Error("flags MFMA_SPLIT and MFMA_F32 exclude each other");
  if (spec->flags & ~(CAST_HIP_FLAG_TILE_SWIZZLE | CAST_HIP_FLAG_TILE_GAUSS | CAST_HIP_FLAG_TILE_XCD |
                      CAST_HIP_FLAG_MFMA_SPLIT | CAST_HIP_FLAG_MFMA_F32))
    throw Error("unknown flags " + std::to_string(spec->flags));
  plan->flags = spec->flags;

  plan->h_mat = new double[2 * K * K];
  std::copy(m.begin(), m.end(), plan->h_mat);

  ensure_device();
  check(hipGetDevice(&plan->device), "hipGetDevice");
  const size_t bytes = 2 * K * K * (fp32 ? 4 : 8);
  plan->d_mat = cache_alloc(bytes);
  plan->mat_bytes = bytes;
  if (fp32) {
    std::vector<float> f(2 * K * K);
    for (size_t i = 0; i < f.size(); ++i)
      f[i] = float(m[i]);
    check(hipMemcpy(plan->d_mat, f.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(matrix)");
  } else {
    check(hipMemcpy(plan->d_mat, m.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(matrix)");
  }
  if (variant != CAST_HIP_VARIANT_MONOMIAL && n_qubits >= 4) {
    // Real matrices of the three-product form: Re M, Im M, Re M + Im M.
    const size_t KK = K * K;
    std::vector<double> m3(3 * KK);
    for (size_t i = 0; i < KK; ++i) {
      m3[i] = m[2 * i];
      m3[KK + i] = m[2 * i + 1];
      m3[2 * KK + i] = m[2 * i] + m[2 * i + 1];
    }
    const size_t b3 = 3 * KK * (fp32 ? 4 : 8);
    plan->d_mat3 = cache_alloc(b3);
    plan->mat3_bytes = b3;
    if (fp32) {
      std::vector<float> f(m3.begin(), m3.end());
      check(hipMemcpy(plan->d_mat3, f.data(), b3, hipMemcpyHostToDevice), "hipMemcpy(m3)");
    } else {
      check(hipMemcpy(plan->d_mat3, m3.data(), b3, hipMemcpyHostToDevice), "hipMemcpy(m3)");
    }
  }
  return plan.release();
}

} // namespace

// Power-of-two size classes from 256 bytes; freed blocks are kept (a plan
// destroyed while other streams run must not wait for the device).
struct DevCache {
  std::mutex mu;
  std::map<size_t, std::vector<void *>> free_blocks;
};
DevCache &dev_cache() {
  static DevCache *c = new DevCache; // intentionally leaked: outlives static teardown
  return *c;
}
size_t size_class(size_t bytes) {
  size_t c = 256;
  while (c < bytes)
    c <<= 1;
  return c;
}

void *cache_alloc(size_t bytes) {
  const size_t c = size_class(bytes);
  DevCache &dc = dev_cache();
  {
    std::lock_guard<std::mutex> lock(dc.mu);
    auto &v = dc.free_blocks[c];
    if (!v.empty()) {
      void *p = v.back();
      v.pop_back();
      return p;
    }
  }
  void *p = nullptr;
  check(hipMalloc(&p, c), "hipMalloc(gate matrix)");
  return p;
}

void cache_release(void *p, size_t bytes) {
  if (!p)
    return;
  DevCache &dc = dev_cache();
  std::lock_guard<std::mutex> lock(dc.mu);
  dc.free_blocks[size_class(bytes)].push_back(p);
}

namespace {

void free_plan(GatePlan *p) {
  if (!p)
    return;
  if (p->d_mat)
    cache_release(p->d_mat, p->mat_bytes);
  if (p->d_mat3)
    cache_release(p->d_mat3, p->mat3_bytes);
  delete[] p->h_mat;
  delete p;
}

} // namespace
} // namespace cast_hip

using namespace cast_hip;

extern "C" {

const char *cast_hip_version(void) { return "cast_hip 0.1.0 (gfx950)"; }

int cast_hip_device_info(cast_hip_device_info_t *out, char *err, size_t len) {
  return guard(err, len, [&] {
    if (!out)
      throw Error("out must not be null");
    ensure_device();
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t p{};
    check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, p.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
    out->n_cu = uint32_t(p.multiProcessorCount);
    out->wavefront_size = uint32_t(p.warpSize);
    out->lds_per_block = uint32_t(p.sharedMemPerBlock);
    out->l2_bytes = uint32_t(p.l2CacheSize);
    out->clock_khz = uint32_t(p.clockRate);
    out->mem_clock_khz = uint32_t(p.memoryClockRate);
    out->mem_bus_width = uint32_t(p.memoryBusWidth);
    out->total_mem = uint64_t(p.totalGlobalMem);
  });
}

int cast_hip_free_memory(uint64_t *out_free, uint64_t *out_total, char *err, size_t len) {
  return guard(err, len, [&] {
    if (!out_free || !out_total)
      throw Error("output pointers must not be null");
    ensure_device();
    size_t f = 0, t = 0;
    check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    *out_free = f;
    *out_total = t;
  });
}

void *cast_hip_stream_create(char *err, size_t len) {
  hipStream_t s = nullptr;
  const int rc = guard(err, len, [&] {
    ensure_device();
    // Default priority, and blocking with respect to the null stream: the
    // synchronous statevector calls (zero, copies, norm) run on the null
    // stream and must stay ordered with gates launched here, as the
    // reference relies on for clone_device (src/cuda/statevector.rs:232-252).
    check(hipStreamCreateWithFlags(&s, hipStreamDefault), "hipStreamCreateWithFlags");
  });
  return rc == 0 ? static_cast<void *>(s) : nullptr;
}

void cast_hip_stream_destroy(void *stream) {
  if (stream)
    (void)hipStreamDestroy(static_cast<hipStream_t>(stream));
}

int cast_hip_stream_sync(void *stream, char *err, size_t len) {
  return guard(err, len, [&] {
    check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
  });
}

void *cast_hip_stream_create_nonblocking(char *err, size_t len) {
  hipStream_t s = nullptr;
  const int rc = guard(err, len, [&] {
    ensure_device();
    check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  });
  return rc == 0 ? static_cast<void *>(s) : nullptr;
}

int cast_hip_event_query(void *event, int *out_done, char *err, size_t len) {
  return guard(err, len, [&] {
    if (!event || !out_done)
      throw Error("null event or output");
    const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
    if (e == hipSuccess) {
      *out_done = 1;
    } else if (e == hipErrorNotReady) {
      (void)hipGetLastError(); // not an error: clear the sticky status
      *out_done = 0;
    } else {
      check(e, "hipEventQuery");
    }
  });
}

void *cast_hip_event_create(char *err, size_t len) {
  hipEvent_t e = nullptr;
  const int rc = guard(err, len, [&] {
    ensure_device();
    check(hipEventCreate(&e), "hipEventCreate");
  });
  return rc == 0 ? static_cast<void *>(e) : nullptr;
}

int cast_hip_event_record(void *event, void *stream, char *err, size_t len) {
  return guard(err, len, [&] {
    if (!event)
      throw Error("null event");
    check(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)),
          "hipEventRecord");
  });
}

int cast_hip_event_synchronize(void *event, char *err, size_t len) {
  return guard(err, len, [&] {
    if (!event)
      throw Error("null event");
    check(hipEventSynchronize(static_cast<hipEvent_t>(event)), "hipEventSynchronize");
  });
}

void cast_hip_event_destroy(void *event) {
  if (event)
    (void)hipEventDestroy(static_cast<hipEvent_t>(event));
}

int cast_hip_event_elapsed_ms(void *start, void *end, float *out_ms, char *err, size_t len) {
  return guard(err, len, [&] {
    if (!start || !end || !out_ms)
      throw Error("null argument");
    check(hipEventElapsedTime(out_ms, static_cast<hipEvent_t>(start),
                              static_cast<hipEvent_t>(end)),
          "hipEventElapsedTime");
  });
}

int cast_hip_memcpy_dtod_async(uint64_t dst, uint64_t src, size_t n_bytes, void *stream,
                               char *err, size_t len) {
  return guard(err, len, [&] {
    if (!dst || !src)
      throw Error("null device pointer");
    check(hipMemcpyAsync(reinterpret_cast<void *>(dst), reinterpret_cast<const void *>(src),
                         n_bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)),
          "hipMemcpyAsync(DtoD)");
  });
}

uint64_t cast_hip_malloc(size_t n_bytes, char *err, size_t len) {
  void *p = nullptr;
  const int rc = guard(err, len, [&] {
    if (n_bytes == 0)
      throw Error("n_bytes must be positive");
    ensure_device();
    check(hipMalloc(&p, n_bytes), "hipMalloc");
  });
  return rc == 0 ? reinterpret_cast<uint64_t>(p) : 0;
}

void cast_hip_free(uint64_t dptr) {
  if (dptr)
    (void)hipFree(reinterpret_cast<void *>(dptr));
}

int cast_hip_memcpy_htod_async(uint64_t dst, const void *src, size_t n_bytes, void *stream,
                               char *err, size_t len) {
  return guard(err, len, [&] {
    if (!dst || (!src && n_bytes))
      throw Error("null pointer");
    check(hipMemcpyAsync(reinterpret_cast<void *>(dst), src, n_bytes, hipMemcpyHostToDevice,
                         static_cast<hipStream_t>(stream)),
          "hipMemcpyAsync(HtoD)");
  });
}

int cast_hip_memcpy_dtoh(void *dst, uint64_t src, size_t n_bytes, void *stream, char *err,
                         size_t len) {
  return guard(err, len, [&] {
    if (!src || (!dst && n_bytes))
      throw Error("null pointer");
    check(hipMemcpyAsync(dst, reinterpret_cast<const void *>(src), n_bytes, hipMemcpyDeviceToHost,
                         static_cast<hipStream_t>(stream)),
          "hipMemcpyAsync(DtoH)");
    check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
  });
}

void *cast_hip_host_alloc(size_t n_bytes, char *err, size_t len) {
  void *p = nullptr;
  const int rc = guard(err, len, [&] {
    if (n_bytes == 0)
      throw Error("n_bytes must be positive");
    ensure_device();
    check(hipHostMalloc(&p, n_bytes, hipHostMallocDefault), "hipHostMalloc");
  });
  return rc == 0 ? p : nullptr;
}

void cast_hip_host_free(void *ptr) {
  if (ptr)
    (void)hipHostFree(ptr);
}

int cast_hip_memcpy_dtoh_async(void *dst, uint64_t src, size_t n_bytes, void *stream, char *err,
                               size_t len) {
  return guard(err, len, [&] {
    if (!src || (!dst && n_bytes))
      throw Error("null pointer");
    check(hipMemcpyAsync(dst, reinterpret_cast<const void *>(src), n_bytes, hipMemcpyDeviceToHost,
                         static_cast<hipStream_t>(stream)),
          "hipMemcpyAsync(DtoH)");
  });
}

uint64_t cast_hip_sv_alloc(size_t n_elements, uint8_t precision, char *err, size_t len) {
  void *p = nullptr;
  const int rc = guard(err, len, [&] {
    check_precision(precision);
    if (n_elements == 0)
      throw Error("n_elements must be positive");
    ensure_device();
    check(hipMalloc(&p, n_elements * scalar_bytes(precision)), "hipMalloc(statevector)");
  });
  return rc == 0 ? reinterpret_cast<uint64_t>(p) : 0;
}

void cast_hip_sv_free(uint64_t dptr) {
  if (dptr)
    (void)hipFree(reinterpret_cast<void *>(dptr));
}

int cast_hip_sv_zero(uint64_t dptr, size_t n_elements, uint8_t precision, char *err, size_t len) {
  return guard(err, len, [&] {
    check_sv_args(dptr, n_elements, precision);
    launch_sv_zero(reinterpret_cast<void *>(dptr), n_elements, precision, nullptr);
    check(hipStreamSynchronize(nullptr), "hipStreamSynchronize(zero)");
  });
}

int cast_hip_sv_zero_async(uint64_t dptr, size_t n_elements, uint8_t precision, void *stream,
                           char *err, size_t len) {
  return guard(err, len, [&] {
    check_sv_args(dptr, n_elements, precision);
    launch_sv_zero(reinterpret_cast<void *>(dptr), n_elements, precision,
                   static_cast<hipStream_t>(stream));
  });
}

int cast_hip_sv_upload(uint64_t dptr, const double *host, size_t n_elements, uint8_t precision,
                       char *err, size_t len) {
  return guard(err, len, [&] {
    check_sv_args(dptr, n_elements, precision);
    if (!host)
      throw Error("null host pointer");
    void *d = reinterpret_cast<void *>(dptr);
    if (precision == 1) {
      check(hipMemcpy(d, host, n_elements * sizeof(double), hipMemcpyHostToDevice),
            "hipMemcpy(HtoD)");
    } else {
      std::vector<float> f(n_elements);
      for (size_t i = 0; i < n_elements; ++i)
        f[i] = float(host[i]);
      check(hipMemcpy(d, f.data(), n_elements * sizeof(float), hipMemcpyHostToDevice),
            "hipMemcpy(HtoD f32)");
    }
  });
}

int cast_hip_sv_download(uint64_t dptr, double *host, size_t n_elements, uint8_t precision,
                         char *err, size_t len) {
  return guard(err, len, [&] {
    check_sv_args(dptr, n_elements, precision);
    if (!host)
      throw Error("null host pointer");
    const void *d = reinterpret_cast<const void *>(dptr);
    if (precision == 1) {
      check(hipMemcpy(host, d, n_elements * sizeof(double), hipMemcpyDeviceToHost),
            "hipMemcpy(DtoH)");
    } else {
      std::vector<float> f(n_elements);
      check(hipMemcpy(f.data(), d, n_elements * sizeof(float), hipMemcpyDeviceToHost),
            "hipMemcpy(DtoH f32)");
      for (size_t i = 0; i < n_elements; ++i)
        host[i] = double(f[i]);
    }
  });
}

int cast_hip_norm_squared(uint64_t dptr, size_t n_elements, uint8_t precision, double *out,
                          char *err, size_t len) {
  return guard(err, len, [&] {
    check_sv_args(dptr, n_elements, precision);
    if (!out)
      throw Error("null output pointer");
    *out = norm_squared(reinterpret_cast<const void *>(dptr), n_elements, precision, nullptr);
  });
}

int cast_hip_scale(uint64_t dptr, size_t n_elements, uint8_t precision, double scale, char *err,
                   size_t len) {
  return guard(err, len, [&] {
    check_sv_args(dptr, n_elements, precision);
    launch_scale(reinterpret_cast<void *>(dptr), n_elements, precision, scale, nullptr);
    check(hipStreamSynchronize(nullptr), "hipStreamSynchronize(scale)");
  });
}

int cast_hip_gate_plan_create(const cast_hip_kernel_gen_spec_t *spec,
                              const cast_complex64_t *matrix, size_t matrix_len,
                              const uint32_t *qubits, size_t n_qubits, void **out_plan,
                              char *err, size_t len) {
  return guard(err, len, [&] {
    if (!out_plan)
      throw Error("out_plan must not be null");
    *out_plan = nullptr;
    *out_plan = build_plan(spec, matrix, matrix_len, qubits, n_qubits);
  });
}

void cast_hip_gate_plan_destroy(void *plan) { free_plan(static_cast<GatePlan *>(plan)); }

int cast_hip_gate_plan_info(const void *plan, cast_hip_gate_plan_info_t *out, char *err,
                            size_t len) {
  return guard(err, len, [&] {
    if (!plan || !out)
      throw Error("null argument");
    const GatePlan *p = static_cast<const GatePlan *>(plan);
    std::memset(out, 0, sizeof(*out));
    out->n_qubits = p->k;
    out->precision = p->precision;
    out->variant = p->variant;
    out->nnz_rows = p->variant == CAST_HIP_VARIANT_MONOMIAL ? p->n_active : (1u << p->k);
    out->scalar_nnz = p->scalar_nnz;
    for (uint32_t b = 0; b < p->k; ++b)
      out->qubits[b] = p->q[b];
  });
}

int cast_hip_gate_plan_family(const void *plan, uint32_t sv_n_qubits, uint32_t *out_variant,
                              char *err, size_t len) {
  return guard(err, len, [&] {
    if (!plan || !out_variant)
      throw Error("null argument");
    const GatePlan *p = static_cast<const GatePlan *>(plan);
    if (sv_n_qubits < p->k || sv_n_qubits > 40)
      throw Error("statevector qubit count out of range for this gate");
    *out_variant = plan_family(*p, sv_n_qubits);
  });
}

int cast_hip_gate_apply(const void *plan, void *stream, uint64_t sv_dptr, uint32_t sv_n_qubits,
                        char *err, size_t len) {
  return guard(err, len, [&] {
    if (!plan)
      throw Error("null plan");
    if (!sv_dptr)
      throw Error("null statevector pointer");
    check_aligned(sv_dptr);
    launch_gate(*static_cast<const GatePlan *>(plan), static_cast<hipStream_t>(stream),
                reinterpret_cast<void *>(sv_dptr), sv_n_qubits);
  });
}

int cast_hip_gate_apply_batch(const void *const *plans, const uint64_t *svs,
                              const uint32_t *sv_n_qubits, size_t n_items, void *stream,
                              char *err, size_t len) {
  return guard(err, len, [&] {
    if (n_items == 0)
      return;
    if (!plans || !svs || !sv_n_qubits)
      throw Error("null argument");
    std::vector<BatchItem> items(n_items);
    const GatePlan *p0 = static_cast<const GatePlan *>(plans[0]);
    if (!p0)
      throw Error("null plan at index 0");
    for (size_t i = 0; i < n_items; ++i) {
      const GatePlan *p = static_cast<const GatePlan *>(plans[i]);
      if (!p)
        throw Error("null plan at index " + std::to_string(i));
      if (p->k != p0->k || p->precision != p0->precision)
        throw Error("batched plans must share k and precision");
      if (!svs[i])
        throw Error("null statevector at index " + std::to_string(i));
      check_aligned(svs[i]);
      const uint32_t n = sv_n_qubits[i];
      if (n < p->k || p->q[p->k - 1] >= n || n > 40)
        throw Error("item " + std::to_string(i) + ": gate does not fit a " +
                    std::to_string(n) + "-qubit statevector");
      BatchItem &it = items[i];
      it.sv = reinterpret_cast<void *>(svs[i]);
      it.mat = p->d_mat;
      it.combos = uint64_t(1) << (n - p->k);
      it.k = uint8_t(p->k);
      for (uint32_t b = 0; b < 6; ++b)
        it.q[b] = uint8_t(b < p->k ? p->q[b] : 0);
    }
    launch_gate_batch(items.data(), n_items, p0->precision, p0->k,
                      static_cast<hipStream_t>(stream));
  });
}

int cast_hip_expect_zz(uint64_t sv_dptr, uint32_t sv_n_qubits, uint8_t precision,
                       const uint32_t *pairs, size_t n_pairs, double *out, void *stream,
                       char *err, size_t len) {
  return guard(err, len, [&] {
    check_precision(precision);
    if (!sv_dptr)
      throw Error("null statevector pointer");
    check_aligned(sv_dptr);
    if (n_pairs > 0 && (!pairs || !out))
      throw Error("null pairs/out");
    if (sv_n_qubits > 40)
      throw Error("statevector qubit count out of range");
    for (size_t p = 0; p < 2 * n_pairs; ++p)
      if (pairs[p] >= sv_n_qubits)
        throw Error("pair qubit " + std::to_string(pairs[p]) + " outside the statevector");
    expect_zz(reinterpret_cast<const void *>(sv_dptr), sv_n_qubits, precision, pairs, n_pairs,
              out, static_cast<hipStream_t>(stream));
  });
}

int cast_hip_traj_run(const cast_hip_traj_launch_t *args, void *stream, char *err, size_t len) {
  return guard(err, len, [&] {
    if (!args)
      throw Error("null trajectory launch descriptor");
    launch_traj(*args, static_cast<hipStream_t>(stream));
  });
}

size_t cast_hip_observe_scratch_bytes(uint32_t n_pairs) { return observe_scratch_bytes(n_pairs); }

int cast_hip_observe_async(uint64_t sv_dptr, uint32_t sv_n_qubits, uint8_t precision,
                           const uint32_t *pairs, size_t n_pairs, uint64_t out, uint64_t scratch,
                           void *stream, char *err, size_t len) {
  return guard(err, len, [&] {
    check_precision(precision);
    if (!sv_dptr || !out || !scratch)
      throw Error("null statevector, output or scratch pointer");
    check_aligned(sv_dptr);
    if (n_pairs > 32)
      throw Error("at most 32 qubit pairs");
    if (n_pairs > 0 && !pairs)
      throw Error("null pairs");
    if (sv_n_qubits < 1 || sv_n_qubits > 40)
      throw Error("statevector qubit count out of range");
    for (size_t p = 0; p < 2 * n_pairs; ++p)
      if (pairs[p] >= sv_n_qubits)
        throw Error("pair qubit " + std::to_string(pairs[p]) + " outside the statevector");
    launch_observe(reinterpret_cast<const void *>(sv_dptr), sv_n_qubits, precision, pairs,
                   n_pairs, reinterpret_cast<double *>(out), reinterpret_cast<double *>(scratch),
                   static_cast<hipStream_t>(stream));
  });
}

int cast_hip_observe_noalloc(uint64_t sv_dptr, uint32_t sv_n_qubits, uint8_t precision,
                             const uint32_t *pairs, size_t n_pairs, uint64_t out, void *stream,
                             char *err, size_t len) {
  return guard(err, len, [&] {
    check_precision(precision);
    if (!sv_dptr || !out)
      throw Error("null statevector or output pointer");
    check_aligned(sv_dptr);
    if (n_pairs > observe_noalloc_max_pairs())
      throw Error("at most " + std::to_string(observe_noalloc_max_pairs()) +
                  " qubit pairs (size of the device scratch)");
    if (n_pairs > 0 && !pairs)
      throw Error("null pairs");
    if (sv_n_qubits < 1 || sv_n_qubits > 40)
      throw Error("statevector qubit count out of range");
    for (size_t p = 0; p < 2 * n_pairs; ++p)
      if (pairs[p] >= sv_n_qubits)
        throw Error("pair qubit " + std::to_string(pairs[p]) + " outside the statevector");
    launch_observe_noalloc(reinterpret_cast<const void *>(sv_dptr), sv_n_qubits, precision, pairs,
                           n_pairs, reinterpret_cast<double *>(out),
                           static_cast<hipStream_t>(stream));
  });
}

int cast_hip_inner_product(uint64_t a, uint8_t precision_a, uint64_t b, uint8_t precision_b,
                           uint32_t sv_n_qubits, double *out_re_im, void *stream, char *err,
                           size_t len) {
  return guard(err, len, [&] {
    check_precision(precision_a);
    check_precision(precision_b);
    if (!a || !b || !out_re_im)
      throw Error("null statevector or output pointer");
    check_aligned(a);
    check_aligned(b);
    if (sv_n_qubits < 1 || sv_n_qubits > 40)
      throw Error("statevector qubit count out of range");
    inner_product(reinterpret_cast<const void *>(a), precision_a, reinterpret_cast<const void *>(b),
                  precision_b, sv_n_qubits, out_re_im, static_cast<hipStream_t>(stream));
  });
}

size_t cast_hip_inner_product_scratch_bytes(void) { return inner_scratch_bytes(); }

int cast_hip_inner_product_async(uint64_t a, uint8_t precision_a, uint64_t b, uint8_t precision_b,
                                 uint32_t sv_n_qubits, uint64_t out, uint64_t scratch,
                                 void *stream, char *err, size_t len) {
  return guard(err, len, [&] {
    check_precision(precision_a);
    check_precision(precision_b);
    if (!a || !b || !out || !scratch)
      throw Error("null statevector, output or scratch pointer");
    check_aligned(a);
    check_aligned(b);
    if (sv_n_qubits < 1 || sv_n_qubits > 40)
      throw Error("statevector qubit count out of range");
    launch_inner(reinterpret_cast<const void *>(a), precision_a, reinterpret_cast<const void *>(b),
                 precision_b, sv_n_qubits, reinterpret_cast<double *>(out),
                 reinterpret_cast<double *>(scratch), static_cast<hipStream_t>(stream));
  });
}

int cast_hip_marginal_probabilities(uint64_t sv_dptr, uint32_t sv_n_qubits, uint8_t precision,
                                    const uint32_t *measured_qubits, size_t n_measured,
                                    double *out, void *stream, char *err, size_t len) {
  return guard(err, len, [&] {
    check_precision(precision);
    if (!sv_dptr || !out)
      throw Error("null statevector or output pointer");
    check_aligned(sv_dptr);
    if (sv_n_qubits < 1 || sv_n_qubits > 40)
      throw Error("statevector qubit count out of range");
    if (n_measured && !measured_qubits)
      throw Error("null measured-qubit list");
    if (n_measured > sv_n_qubits)
      throw Error("more measured qubits than the statevector has");
    marginal_probabilities(reinterpret_cast<const void *>(sv_dptr), sv_n_qubits, precision,
                           measured_qubits, uint32_t(n_measured), out,
                           static_cast<hipStream_t>(stream));
  });
}

int cast_hip_dm_populations(uint64_t sv_dptr, uint32_t n_physical, uint8_t precision, double *out,
                            void *stream, char *err, size_t len) {
  return guard(err, len, [&] {
    check_precision(precision);
    if (!sv_dptr || !out)
      throw Error("null statevector or output pointer");
    check_aligned(sv_dptr);
    if (n_physical < 1 || n_physical > 20)
      throw Error("density-matrix qubit count out of range (1..20)");
    dm_populations(reinterpret_cast<const void *>(sv_dptr), n_physical, precision, out,
                   static_cast<hipStream_t>(stream));
  });
}

} // extern "C"
