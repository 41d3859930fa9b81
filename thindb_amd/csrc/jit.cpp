// jit.cpp — plan-specialised scan kernels, compiled at run time with hipRTC.
//
// The generic kernel interprets the plan (column kinds, predicate terms, key
// parts, aggregate kinds) on every strip; that costs about half of its
// bandwidth.  Here a plan's *structure* becomes straight-line code: each
// column is loaded with its kind fixed at compile time, each predicate and
// aggregate is one inlined template instance, and only values (literals,
// key bases, offsets, pointers) still come from the kernel argument.  The
// code object is cached per (device, signature), so repeated queries of one
// shape (a serving loop, a batch) compile once.  When hipRTC is unavailable
// or THINDB_GPU_JIT=0, the runtime uses the generic kernel.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "jit_headers.inc"  // kJitHeaderSrc / kJitHeaderName (generated by the build)

namespace tdbg {

namespace {

struct JitKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  bool failed = false;
};

std::mutex g_mu;
std::unordered_map<std::string, JitKernel> g_cache;
std::string g_last_log;

const char* kind_name(uint32_t k) {
  switch (k) {
    case K_U8: return "K_U8";
    case K_I8: return "K_I8";
    case K_I16: return "K_I16";
    case K_I32: return "K_I32";
    case K_I64: return "K_I64";
    case K_F32: return "K_F32";
    default: return "K_F64";
  }
}
const char* pred_name(uint32_t k) {
  switch (k) {
    case P_RANGE_I: return "P_RANGE_I";
    case P_RANGE_F: return "P_RANGE_F";
    case P_NEQ_I: return "P_NEQ_I";
    default: return "P_NEQ_F";
  }
}
const char* agg_name(uint32_t k) {
  switch (k) {
    case A_COUNT_COL: return "A_COUNT_COL";
    case A_SUM_I: return "A_SUM_I";
    case A_SUM_F: return "A_SUM_F";
    case A_MIN_I: return "A_MIN_I";
    case A_MAX_I: return "A_MAX_I";
    case A_MIN_F: return "A_MIN_F";
    default: return "A_MAX_F";
  }
}
const char* lmode_name(uint32_t m) {
  return m == LDS_DIRECT ? "LDS_DIRECT" : (m == LDS_HASH ? "LDS_HASH" : "LDS_NONE");
}

bool jit_enabled() {
  const char* e = std::getenv("THINDB_GPU_JIT");
  return !(e && e[0] == '0');
}

}  // namespace

// The structure of a plan: everything the generated code fixes.
std::string jit_signature(const ScanParams& p) {
  std::ostringstream s;
  s << "m" << p.lds_mode << "k" << p.lds_key64 << "c";
  for (uint32_t i = 0; i < p.n_cols; i++) s << p.cols[i].kind << (p.cols[i].valid ? "v" : "") << ",";
  s << "p";
  for (uint32_t i = 0; i < p.n_preds; i++) s << p.preds[i].kind << ":" << p.preds[i].col << ",";
  s << "k";
  for (uint32_t i = 0; i < p.n_parts; i++) s << p.parts[i].col << (p.parts[i].nullable ? "n" : "") << ",";
  s << "f";
  for (uint32_t i = 0; i < p.n_filt; i++)
    s << p.filt[i].kind << ":" << p.filt[i].arity << ":" << p.filt[i].pd.kind << ":" << p.filt[i].pd.col << ",";
  s << "a";
  for (uint32_t i = 0; i < p.n_aggs; i++)
    s << p.aggs[i].kind << ":" << p.aggs[i].col << (p.aggs[i].nullable ? "n" : "") << (p.aggs[i].split ? "s" : "")
      << ",";
  return s.str();
}

// The strip of a plan's structure as straight-line code: struct `name`
// with load / run / process, the interface scan_range and scan_range_pf use.
std::string strip_struct(const ScanParams& p, const std::string& name) {
  std::ostringstream s;
  s << "struct " << name << " {\n"
       "  struct Data {\n";
  for (uint32_t i = 0; i < p.n_cols; i++) s << "    Raw x" << i << ";\n";
  s << "  };\n"
       "  static __device__ __forceinline__ void load(const ScanParams& P, uint64_t r, Data& d) {\n";
  for (uint32_t i = 0; i < p.n_cols; i++)
    s << "    d.x" << i << " = ld<" << kind_name(p.cols[i].kind) << ">(P.cols[" << i << "], r);\n";
  s << "  }\n"
       "  template <uint32_t LMODE>\n"
       "  static __device__ __forceinline__ void run(const ScanParams& P, uint8_t* lds, uint64_t r, uint32_t cnt,\n"
       "                                             uint64_t part_lo) {\n"
       "    Data d;\n"
       "    load(P, r, d);\n"
       "    process<LMODE>(P, lds, r, cnt, part_lo, d);\n"
       "  }\n"
       "  template <uint32_t LMODE>\n"
       "  static __device__ __forceinline__ void process(const ScanParams& P, uint8_t* lds, uint64_t r, uint32_t cnt,\n"
       "                                                 uint64_t part_lo, const Data& d) {\n";
  for (uint32_t i = 0; i < p.n_cols; i++) s << "    const Raw& x" << i << " = d.x" << i << ";\n";
  s << "    uint32_t mask = (1u << cnt) - 1u;\n";
  for (uint32_t i = 0; i < p.n_preds; i++) {
    const PredDesc& d = p.preds[i];
    s << "    mask &= pred_k<" << pred_name(d.kind) << ", " << kind_name(p.cols[d.col].kind) << ">(P.preds[" << i
      << "], x" << d.col << ")";
    if (p.cols[d.col].valid) s << " & valid4(P.cols[" << d.col << "], r)";
    s << ";\n";
  }
  if (p.n_filt) {
    // the WHERE tree, its postfix order unrolled into named strip masks
    std::vector<std::string> stk;
    for (uint32_t i = 0; i < p.n_filt; i++) {
      const FilterNode& f = p.filt[i];
      const std::string v = "f" + std::to_string(i);
      const bool valid = f.kind <= H_IS_NOT_NULL && p.cols[f.pd.col].valid;
      const std::string vcol = "valid4(P.cols[" + std::to_string(f.pd.col) + "], r)";
      std::string e;
      switch (f.kind) {
        case H_CMP:
          e = std::string("pred_k<") + pred_name(f.pd.kind) + ", " + kind_name(p.cols[f.pd.col].kind) + ">(P.filt[" +
              std::to_string(i) + "].pd, x" + std::to_string(f.pd.col) + ")" + (valid ? " & " + vcol : "");
          break;
        case H_IS_NULL: e = valid ? "(~" + vcol + " & 0xfu)" : "0u"; break;
        case H_IS_NOT_NULL: e = valid ? vcol : "0xfu"; break;
        case H_NOT:
          e = "(~" + stk.back() + " & 0xfu)";
          stk.pop_back();
          break;
        default: {
          const char* op = f.kind == H_AND ? " & " : " | ";
          e = f.kind == H_AND ? "0xfu" : "0u";
          for (uint32_t a = 0; a < f.arity; a++) {
            e += op + stk.back();
            stk.pop_back();
          }
          break;
        }
      }
      s << "    const uint32_t " << v << " = " << e << ";\n";
      stk.push_back(v);
    }
    s << "    mask &= " << stk.back() << ";\n";
  }
  if (p.n_preds || p.n_filt) s << "    if (!mask) return;\n";
  s << "    uint64_t key[kVec];\n";
  for (uint32_t i = 0; i < p.n_parts; i++) {
    const PartDesc& d = p.parts[i];
    s << "    { int64_t v[kVec]; dec_i<" << kind_name(p.cols[d.col].kind) << ">(x" << d.col << ", v); key_part<"
      << (i == 0 ? "true" : "false") << ">(P, P.parts[" << i << "], v, key, mask, ";
    if (d.nullable) s << "valid4(P.cols[" << d.col << "], r)";
    else s << "0xfu";
    s << "); }\n";
  }
  s << "    if (!mask) return;\n"
       "    int slot[kVec];\n"
       "    uint8_t* rec[kVec];\n"
       "    mask = resolve<LMODE>(P, lds, key, mask, part_lo, slot, rec);\n";
  for (uint32_t i = 0; i < p.n_aggs; i++) {
    const AggDesc& d = p.aggs[i];
    s << "    { const uint32_t m = mask";
    if (p.cols[d.col].valid) s << " & valid4(P.cols[" << d.col << "], r)";
    s << "; if (m) agg_k<" << agg_name(d.kind) << ", " << kind_name(p.cols[d.col].kind) << ", "
      << (d.nullable ? "true" : "false") << ", " << (d.split ? "true" : "false") << ">(P, P.aggs[" << i << "], x"
      << d.col << ", m, slot, rec, lds); }\n";
  }
  s << "  }\n"
       "};\n";
  return s.str();
}

std::string jit_source(const ScanParams& p, bool prefetch) {
  std::ostringstream s;
  s << "#include \"scan_common.h\"\n"
       "namespace tdbg {\n"
    << strip_struct(p, "JitStrip")
    << "}  // namespace tdbg\n"
       "extern \"C\" __global__ void __launch_bounds__(1024) tdbg_scan_jit(const tdbg::ScanParams P) {\n"
       "  extern __shared__ __align__(16) uint8_t lds[];\n"
       "  tdbg::"
    << (prefetch ? "scan_body_pf" : "scan_body") << "<tdbg::" << lmode_name(p.lds_mode)
    << ", tdbg::JitStrip>(P, lds);\n"
       "}\n";
  return s.str();
}

// A batch kernel for a batch's scan plans: batch_device.h's loop with one
// generated strip per distinct plan structure (BatchQuery::jit_sig picks
// it; ~0u and radix plans take the generic paths).
std::string batch_jit_source(const std::vector<ScanParams>& sigs, bool generic) {
  std::ostringstream s;
  s << "#include \"batch_device.h\"\n";
  if (generic) s << "#include \"strip_generic.h\"\n";
  s << "namespace tdbg {\n";
  for (size_t i = 0; i < sigs.size(); i++) s << strip_struct(sigs[i], "JitStrip" + std::to_string(i));
  s << "struct JitScan {\n"
       "  static __device__ __forceinline__ void run(const BatchQuery& B, const ScanParams& P, uint8_t* lds,\n"
       "                                             const BatchItem& it) {\n"
       "    const uint64_t part_lo = (uint64_t)it.part * P.lds_cap;\n"
       "    (void)part_lo;\n"
       "    switch (B.jit_sig) {\n";
  for (size_t i = 0; i < sigs.size(); i++) {
    const uint32_t m = sigs[i].lds_mode;
    s << "      case " << i << ": scan_range<" << lmode_name(m) << ", JitStrip" << i << ">(P, lds, it.a0, it.a1, "
      << (m == LDS_DIRECT ? "part_lo" : "0") << "); break;\n";
  }
  s << "      default:\n";
  if (generic)
    s << "        if (P.lds_mode == LDS_DIRECT) scan_range<LDS_DIRECT, GenericStrip<false>>(P, lds, it.a0, it.a1, part_lo);\n"
         "        else if (P.lds_mode == LDS_HASH) scan_range<LDS_HASH, GenericStrip<false>>(P, lds, it.a0, it.a1, 0);\n"
         "        else scan_range<LDS_NONE, GenericStrip<false>>(P, lds, it.a0, it.a1, 0);\n";
  s << "        break;\n"
       "    }\n"
       "  }\n"
       "};\n"
       "}  // namespace tdbg\n"
       "extern \"C\" __global__ void __launch_bounds__(tdbg::kBatchThreads, 4) tdbg_batch_jit(\n"
       "    const tdbg::BatchQuery* __restrict__ Q, const tdbg::BatchItem* __restrict__ items,\n"
       "    const tdbg::BatchArgs* __restrict__ Ap) {\n"
       "  tdbg::batch_loop<tdbg::JitScan>(Q, items, Ap);\n"
       "}\n";
  return s.str();
}

// The specialised kernel loads the next strip while it aggregates the
// current one (filtered 200M-row query: 1.149 -> 1.085 ms);
// THINDB_GPU_JIT_PREFETCH=0 turns that off.
bool jit_prefetch() {
  const char* e = std::getenv("THINDB_GPU_JIT_PREFETCH");
  return !(e && e[0] == '0');
}

static bool compile(const std::string& src, std::vector<char>& code, std::string& log) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "tdbg_jit.hip", kJitHeaders, kJitHeaderSrc, kJitHeaderName) !=
      HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return false;
  }
  int dev = 0;
  hipGetDevice(&dev);
  hipDeviceProp_t prop;
  hipGetDeviceProperties(&prop, dev);
  std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
  const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-munsafe-fp-atomics",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, 8, opts);
  size_t log_size = 0;
  hiprtcGetProgramLogSize(prog, &log_size);
  if (log_size > 1) {
    std::vector<char> l(log_size + 1, 0);
    hiprtcGetProgramLog(prog, l.data());
    log = l.data();
  }
  bool ok = rc == HIPRTC_SUCCESS;
  if (ok) {
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.resize(n);
    ok = n > 0 && hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
  }
  hiprtcDestroyProgram(&prog);
  return ok;
}

// Launches the plan-specialised kernel; returns hipErrorNotSupported when
// run-time compilation is unavailable so the caller takes the generic one.
hipError_t launch_scan_jit(const ScanParams& p, uint32_t grid, uint32_t block, hipStream_t s, bool* used,
                           bool cached_only) {
  *used = false;
  if (!jit_enabled() || p.n_cols > 8) return hipErrorNotSupported;
  int dev = 0;
  hipGetDevice(&dev);
  const bool pf = jit_prefetch();
  const std::string key = std::to_string(dev) + "|" + (pf ? "pf|" : "") + jit_signature(p);
  hipFunction_t fn = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (cached_only) {
      auto it = g_cache.find(key);
      if (it == g_cache.end() || !it->second.fn) return hipErrorNotSupported;
    }
    JitKernel& k = g_cache[key];
    if (k.failed) return hipErrorNotSupported;
    if (!k.fn) {
      std::vector<char> code;
      std::string log;
      if (!compile(jit_source(p, pf), code, log)) {
        k.failed = true;
        g_last_log = log;
        if (std::getenv("THINDB_GPU_JIT_VERBOSE")) fprintf(stderr, "[tdbg jit] compile failed: %s\n", log.c_str());
        return hipErrorNotSupported;
      }
      if (hipModuleLoadData(&k.mod, code.data()) != hipSuccess ||
          hipModuleGetFunction(&k.fn, k.mod, "tdbg_scan_jit") != hipSuccess) {
        (void)hipGetLastError();
        k.failed = true;
        return hipErrorNotSupported;
      }
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k.fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                160 * 1024);
      (void)hipGetLastError();
    }
    fn = k.fn;
  }
  ScanParams arg = p;
  void* args[] = {&arg};
  const size_t lds = p.lds_mode == LDS_NONE ? 0 : p.lds_bytes;
  hipError_t e = hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, (unsigned)lds, s, args, nullptr);
  if (e == hipSuccess) *used = true;
  return e;
}

// The batch kernel for the scan-plan structures `sigs` (compiled once per
// device and set of structures); nullptr when run-time compilation is off or
// fails, and the batch takes the ahead-of-time kernel.
hipFunction_t batch_jit_kernel(const std::vector<ScanParams>& sigs, bool generic) {
  if (!jit_enabled() || sigs.empty()) return nullptr;
  const char* e = std::getenv("THINDB_GPU_BATCH_JIT");
  if (e && e[0] == '0') return nullptr;
  int dev = 0;
  hipGetDevice(&dev);
  std::string key = std::to_string(dev) + "|batch|" + (generic ? "g|" : "");
  for (const ScanParams& p : sigs) key += jit_signature(p) + "#";
  std::lock_guard<std::mutex> lk(g_mu);
  JitKernel& k = g_cache[key];
  if (k.failed) return nullptr;
  if (!k.fn) {
    std::vector<char> code;
    std::string log;
    if (!compile(batch_jit_source(sigs, generic), code, log)) {
      k.failed = true;
      g_last_log = log;
      if (std::getenv("THINDB_GPU_JIT_VERBOSE")) fprintf(stderr, "[tdbg jit] batch compile failed: %s\n", log.c_str());
      return nullptr;
    }
    if (hipModuleLoadData(&k.mod, code.data()) != hipSuccess ||
        hipModuleGetFunction(&k.fn, k.mod, "tdbg_batch_jit") != hipSuccess) {
      (void)hipGetLastError();
      k.failed = true;
      k.fn = nullptr;
      return nullptr;
    }
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k.fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024 - 1024);
    (void)hipGetLastError();
  }
  return k.fn;
}

hipError_t launch_batch_jit(hipFunction_t fn, const BatchQuery* Q, const BatchItem* items, const BatchArgs* args,
                            uint32_t grid, uint32_t lds_bytes, hipStream_t s) {
  void* kargs[] = {&Q, &items, &args};
  return hipModuleLaunchKernel(fn, grid, 1, 1, kBatchThreads, 1, 1, lds_bytes, s, kargs, nullptr);
}

const char* jit_last_log() { return g_last_log.c_str(); }

}  // namespace tdbg
