// Context, device memory, event timing and host<->device helpers.
#include "jues_internal.h"

#include <atomic>

namespace jues {

static std::atomic<size_t> g_bytes{0};
size_t device_bytes_in_use() { return g_bytes.load(); }

static thread_local bool tl_no_alloc = false;
NoAllocScope::NoAllocScope() : prev(tl_no_alloc) { tl_no_alloc = true; }
NoAllocScope::~NoAllocScope() { tl_no_alloc = prev; }
bool no_alloc_active() { return tl_no_alloc; }

void DevBuf::alloc(size_t nbytes) {
  release();
  if (nbytes == 0) return;
  if (tl_no_alloc) throw Error(JUES_BAD_ARG, "device allocation inside an allocation-free entry point");
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, nbytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(e == hipErrorOutOfMemory ? JUES_OOM : JUES_HIP_ERROR,
                "hipMalloc(" + std::to_string(nbytes) + " bytes): " + hipGetErrorString(e));
  }
  p = q;
  bytes = nbytes;
  g_bytes += nbytes;
}

void DevBuf::release() {
  if (p) {
    (void)hipFree(p);
    g_bytes -= bytes;
  }
  p = nullptr;
  bytes = 0;
}

namespace {
struct ScratchPool {
  DevBuf slot[20];
};
std::map<const jues_ctx*, ScratchPool>& pools() {
  static std::map<const jues_ctx*, ScratchPool> m;
  return m;
}
}  // namespace

void* scratch(jues_ctx* ctx, int slot, size_t bytes) {
  if (tl_no_alloc) throw Error(JUES_BAD_ARG, "context scratch inside an allocation-free entry point");
  DevBuf& b = pools()[ctx].slot[slot];
  if (b.bytes < bytes) {
    if (b.p) JUES_HIP(hipStreamSynchronize(ctx->stream));
    b.alloc(bytes + bytes / 8 + 256);
  }
  return b.p;
}

void release_scratch(jues_ctx* ctx) { pools().erase(ctx); }

void h2d(jues_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  JUES_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  JUES_HIP(hipStreamSynchronize(ctx->stream));
}

void d2h(jues_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  JUES_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  JUES_HIP(hipStreamSynchronize(ctx->stream));
}

void sync(jues_ctx* ctx) { JUES_HIP(hipStreamSynchronize(ctx->stream)); }

KTimer::KTimer(jues_ctx* c, const char* name, double flops, double bytes) : ctx(c), on(c->timing) {
  if (!on) return;
  hipEvent_t a, b;
  if (ctx->event_pool.size() >= 2) {
    a = ctx->event_pool.back(); ctx->event_pool.pop_back();
    b = ctx->event_pool.back(); ctx->event_pool.pop_back();
  } else {
    JUES_HIP(hipEventCreate(&a));
    JUES_HIP(hipEventCreate(&b));
  }
  JUES_HIP(hipEventRecord(a, ctx->stream));
  slot = ctx->pending.size();
  ctx->pending.push_back({name, a, b, flops, bytes});
}

KTimer::~KTimer() {
  if (!on) return;
  (void)hipEventRecord(ctx->pending[slot].b, ctx->stream);
  if (ctx->pending.size() > 4096) {
    try { flush_timers(ctx); } catch (...) {}
  }
}

void flush_timers(jues_ctx* ctx) {
  if (ctx->pending.empty()) return;
  JUES_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->pending) {
    float ms = 0.f;
    JUES_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    KStat& s = ctx->kstats[p.name];
    s.ms += ms;
    s.calls += 1;
    s.flops += p.flops;
    s.bytes += p.bytes;
    ctx->event_pool.push_back(p.a);
    ctx->event_pool.push_back(p.b);
  }
  ctx->pending.clear();
}

}  // namespace jues

jues_ctx::jues_ctx() {}
jues_ctx::~jues_ctx() {}
