// Kernel definition and launch layer.
//
// Every kernel of the library is written as a device body whose first three
// parameters are a launch-bounds tag and its workgroup index and grid size
// (named lw_bid_ and lw_gdim_; the names blockIdx and gridDim are mapped to
// them in the kernel sources). Two global wrappers call a body:
// - k_one: one case, the hardware's workgroup index and grid size;
// - k_bat: several cases in one launch. The launch's workgroups are cut into
//   consecutive ranges, one per case (BItem); a workgroup finds its range,
//   takes its index within it and its case's grid size, reads the case's
//   arguments from a device record, and runs the same body. Each range starts
//   at a multiple of 8 workgroups, so the body's placement of workgroups on
//   the 8 XCDs is the same as in a launch of its own.
// A body runs the same arithmetic either way: a case gives the same bits in
// a batched launch as alone.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace lw {

struct Blk {
  unsigned x, y, z;
};
// launch bounds (threads per workgroup) and waves per EU (0: no attribute)
template <int LB, int WPE>
struct Lb {};

// one case's workgroups in a batched launch
struct BItem {
  unsigned blk0;        // its first workgroup (a multiple of 8)
  unsigned gx, gy, gz;  // its grid
  const char *args;     // its argument record (device memory)
};

// Argument records: an argument of at most 16 bytes (and alignment <= 8)
// sits in the record itself, in an 8-byte aligned slot; a larger one (the
// geometry, field and multigrid tables) is stored once per flush and the
// slot holds its device address.
template <typename T>
constexpr bool arg_by_ref() {
  return sizeof(T) > 16 || alignof(T) > 8;
}
template <typename T>
constexpr unsigned arg_slot() {
  return arg_by_ref<T>() ? 8u : (unsigned)((sizeof(T) + 7) & ~(size_t)7);
}
template <typename... A>
struct ArgLayout {
  static constexpr unsigned n = sizeof...(A);
  static constexpr unsigned size[n + 1] = {arg_slot<A>()..., 0u};
  static constexpr unsigned off(unsigned i) {
    unsigned o = 0;
    for (unsigned q = 0; q < i; ++q) o += size[q];
    return o;
  }
  static constexpr unsigned bytes = off(n);
};

// argument records are read through the constant address space: the
// workgroup-uniform addresses then become scalar loads. A body parameter of
// reference type binds to the stored value itself (no private copy).
template <typename A>
__device__ __forceinline__ decltype(auto) arg_load(const char *p) {
  typedef std::decay_t<A> T;
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) char *CP;
  CP q = (CP)p;
  if constexpr (arg_by_ref<T>()) {
    typedef const __attribute__((address_space(4))) T *TP;
    TP b = *(const __attribute__((address_space(4))) TP *)q;
    if constexpr (std::is_reference<A>::value) {
      return (const T &)*(const T *)b;
    } else {
      T v = *(const T *)b;
      return v;
    }
  } else {
    T v = *(const __attribute__((address_space(4))) T *)q;
    return v;
  }
#else
  (void)p;
  if constexpr (std::is_reference<A>::value) {
    return (const T &)*(const T *)nullptr;
  } else {
    return T{};
  }
#endif
}

template <auto BODY, int LB, int WPE, typename... A, size_t... I>
__device__ __forceinline__ void bat_call(const char *p, Blk b, Blk g, std::index_sequence<I...>) {
  using L = ArgLayout<std::decay_t<A>...>;
  BODY(Lb<LB, WPE>{}, b, g, arg_load<A>(p + L::off(I))...);
}

template <int LB, int WPE, auto BODY, typename... A>
__device__ __forceinline__ void bat_run(const BItem *__restrict__ items, int n) {
  const unsigned b = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const BItem it = items[lo];
  const unsigned l = b - it.blk0;
  if (l >= it.gx * it.gy * it.gz) return;  // the padding of a range to 8 workgroups
  bat_call<BODY, LB, WPE, A...>(it.args, Blk{l % it.gx, (l / it.gx) % it.gy, l / (it.gx * it.gy)},
                                Blk{it.gx, it.gy, it.gz}, std::index_sequence_for<A...>{});
}

// A are the body's parameter types; the kernels take them by value
template <auto BODY, int LB, typename... A>
__global__ __launch_bounds__(LB) void k_one(std::decay_t<A>... a) {
  BODY(Lb<LB, 0>{}, Blk{blockIdx.x, blockIdx.y, blockIdx.z}, Blk{gridDim.x, gridDim.y, gridDim.z}, a...);
}
template <auto BODY, int LB, int WPE, typename... A>
__global__ __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_one_w(std::decay_t<A>... a) {
  BODY(Lb<LB, WPE>{}, Blk{blockIdx.x, blockIdx.y, blockIdx.z}, Blk{gridDim.x, gridDim.y, gridDim.z}, a...);
}
template <auto BODY, int LB, typename... A>
__global__ __launch_bounds__(LB) void k_bat(const BItem *__restrict__ items, int n) {
  bat_run<LB, 0, BODY, A...>(items, n);
}
template <auto BODY, int LB, int WPE, typename... A>
__global__ __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_bat_w(const BItem *__restrict__ items,
                                                                                       int n) {
  bat_run<LB, WPE, BODY, A...>(items, n);
}

// a kernel as its two wrappers; A are the kernel's (by-value) parameter types
template <typename... A>
struct KFn {
  const void *one_ = nullptr, *bat_ = nullptr;
  typedef void (*fn_t)(A...);
  fn_t one() const { return (fn_t)one_; }
  explicit operator bool() const { return one_ != nullptr; }
};
// the dynamic LDS limit of both wrappers of a kernel
template <typename... A>
inline hipError_t set_max_lds(const KFn<A...> &k, int bytes) {
  hipError_t e = hipFuncSetAttribute(k.one_, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess) e = hipFuncSetAttribute(k.bat_, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  return e;
}
template <typename F>
struct KMake;
template <int LB, int WPE, typename... A>
struct KMake<void (*)(Lb<LB, WPE>, Blk, Blk, A...)> {
  template <auto BODY>
  static KFn<std::decay_t<A>...> make() {
    if constexpr (WPE > 0)
      return KFn<std::decay_t<A>...>{(const void *)k_one_w<BODY, LB, WPE, A...>,
                                     (const void *)k_bat_w<BODY, LB, WPE, A...>};
    else
      return KFn<std::decay_t<A>...>{(const void *)k_one<BODY, LB, A...>, (const void *)k_bat<BODY, LB, A...>};
  }
};
template <auto BODY>
inline auto kfn() {
  return KMake<decltype(BODY)>::template make<BODY>();
}

// ---- recording of launches (batched stepping) ----
// While a thread records (g_rec set), the launch macros and the step's
// stream operations append to its recorder instead of reaching the device;
// the batch merges the recorders of its cases and issues them.
struct RecOp {
  int kind;  // 0 launch, 1 read-back (device src to host dst at the flush's end)
  const void *bat;
  unsigned bx, by, bz, gx, gy, gz, shm;
  uint32_t arg_off, arg_len, fix_off, fix_n;
  void *dst;
  const void *src;
  size_t bytes;
  const char *name;
};
struct RecFix {
  uint32_t slot;  // byte offset of the slot in the recorder's argument records
  uint32_t val;   // index of the stored value
};
struct Recorder {
  std::vector<RecOp> ops;
  std::vector<char> args;
  std::vector<RecFix> fix;
  std::vector<char> vals;  // stored by-reference values, 16-byte aligned
  std::vector<uint32_t> val_off, val_len;
  uint32_t put_val(const void *p, size_t n) {
    // the same table is passed to many launches: reuse a recent copy
    const size_t nv = val_off.size();
    for (size_t q = nv, c = 0; q > 0 && c < 24; --q, ++c)
      if (val_len[q - 1] == n && memcmp(vals.data() + val_off[q - 1], p, n) == 0) return (uint32_t)(q - 1);
    const size_t o = (vals.size() + 15) & ~(size_t)15;
    vals.resize(o + n);
    memcpy(vals.data() + o, p, n);
    val_off.push_back((uint32_t)o);
    val_len.push_back((uint32_t)n);
    return (uint32_t)nv;
  }
  void clear() {
    ops.clear();
    args.clear();
    fix.clear();
    vals.clear();
    val_off.clear();
    val_len.clear();
  }
};
extern thread_local Recorder *g_rec;

template <typename... A, size_t... I>
inline void rec_store(Recorder &r, uint32_t base, const std::tuple<A...> &t, std::index_sequence<I...>) {
  using L = ArgLayout<A...>;
  (
      [&] {
        typedef std::tuple_element_t<I, std::tuple<A...>> T;
        static_assert(std::is_trivially_copyable<T>::value, "kernel arguments are copied as bytes");
        const uint32_t slot = base + L::off(I);
        if constexpr (arg_by_ref<T>()) {
          const uint32_t v = r.put_val(&std::get<I>(t), sizeof(T));
          r.fix.push_back(RecFix{slot, v});
        } else {
          memcpy(r.args.data() + slot, &std::get<I>(t), sizeof(T));
        }
      }(),
      ...);
}

template <typename... A, typename... X>
inline void rec_launch(const KFn<A...> &k, const char *name, dim3 grid, dim3 block, size_t shm, X &&...x) {
  static_assert(sizeof...(A) == sizeof...(X), "argument count of the kernel");
  using L = ArgLayout<A...>;
  Recorder &r = *g_rec;
  const std::tuple<A...> t(static_cast<A>(std::forward<X>(x))...);
  RecOp op{};
  op.kind = 0;
  op.bat = k.bat_;
  op.bx = block.x;
  op.by = block.y;
  op.bz = block.z;
  op.gx = grid.x;
  op.gy = grid.y;
  op.gz = grid.z;
  op.shm = (unsigned)shm;
  op.arg_off = (uint32_t)r.args.size();
  op.arg_len = L::bytes;
  op.fix_off = (uint32_t)r.fix.size();
  op.name = name;
  r.args.resize(r.args.size() + L::bytes, 0);
  rec_store(r, op.arg_off, t, std::index_sequence_for<A...>{});
  op.fix_n = (uint32_t)r.fix.size() - op.fix_off;
  if (op.gx == 0 || op.gy == 0 || op.gz == 0) return;  // an empty launch does nothing
  r.ops.push_back(op);
}

// batched stepping (lw_batch.hip): recorded zero fills and read-backs, the
// rendezvous of a recording thread, and the batch itself
void rec_zero(void *p, size_t bytes);
void rec_read(void *host, const void *dev, size_t bytes);
void batch_sync();
void batch_gate();
struct Batch;
Batch *batch_new(int dev, int n, int (*run)(void *, int, int64_t), void *ctx);
int batch_run(Batch *b, int64_t nsteps, int *rc);
void batch_counts(const Batch *b, int64_t out[6]);
// the message of the device error that ended the batch (empty if none)
const std::string &batch_error(const Batch *b);

void batch_free(Batch *b);

}  // namespace lw

// a kernel body's leading parameters
#define LW_KERNEL __device__ __forceinline__
#define LW_KP(LB, WPE) ::lw::Lb<LB, WPE>, const ::lw::Blk lw_bid_, const ::lw::Blk lw_gdim_
// for device helpers that read the workgroup index or grid size
#define LW_KB const ::lw::Blk lw_bid_, const ::lw::Blk lw_gdim_
#define LW_KBA lw_bid_, lw_gdim_
#define LW_KFN(K) (::lw::kfn<K>())
