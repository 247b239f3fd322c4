// CPU emulation of the small HIP subset the engine's kernels use, for host-side
// tests of kernel logic (built with g++ and AddressSanitizer, never with hipcc).
//
//  * A launch hands its blocks to EMU_WORKERS OS threads (default 1; several
//    workers run blocks concurrently, as persistent kernels need).  Inside a
//    block every GPU thread is a user-level fiber with its own stack; a worker
//    runs its fibers round robin, switching only at __syncthreads() and at
//    cross-lane operations (__shfl*, readlane), which are barriers over the
//    block or the wavefront.  Fibers that return leave both barriers.
//  * __shared__ variables become thread_local statics: one copy per worker,
//    i.e. per running block.  Dynamic LDS is a buffer of exactly the launch's
//    size.  LDS and device memory start filled with a byte pattern so that reads
//    of unwritten memory show up; AddressSanitizer reports any access outside an
//    allocation.
#pragma once
#include <math.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __align__(n) alignas(n)

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

inline thread_local dim3 threadIdx;
inline thread_local dim3 blockIdx;
inline dim3 blockDim;
inline dim3 gridDim;

// x86-64 System V context switch: saves the callee-saved registers on the current
// stack, stores the stack pointer in *from, loads `to` and restores from there.
extern "C" void emu_ctx_switch(void** from, void* to);
asm(R"(
.text
.weak emu_ctx_switch
.type emu_ctx_switch,@function
emu_ctx_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size emu_ctx_switch,.-emu_ctx_switch
)");

namespace emu {

inline constexpr unsigned char kFill = 0x7f;
inline size_t stack_bytes() {
    static const size_t n = [] {
        const char* e = std::getenv("EMU_STACK_KB");
        return (size_t)(e ? std::max(16, std::atoi(e)) : 512) * 1024;
    }();
    return n;
}

struct Fiber {
    void* sp = nullptr;
    std::unique_ptr<unsigned char[]> stack;   // uninitialised; reused across blocks
    size_t stack_size = 0;
    dim3 tid;
    unsigned flat = 0;
    bool done = false;
    bool waiting = false;
    void* fake_stack = nullptr;
};

struct Barrier {
    int n = 0, arrived = 0;
    std::vector<int> waiters;
};

struct Block {
    std::vector<Fiber> fib;
    Barrier bar;
    std::vector<Barrier> wave;
    std::vector<uint64_t> xchg;
    std::vector<unsigned char> lds;
    std::function<void()> body;
    void* sched_sp = nullptr;
    int cur = -1;
    const void* sched_bottom = nullptr;
    size_t sched_size = 0;
};

inline thread_local Block* tl_block = nullptr;

inline void switch_to_sched() {
    Block* b = tl_block;
    Fiber& f = b->fib[b->cur];
#if EMU_ASAN
    __sanitizer_start_switch_fiber(&f.fake_stack, b->sched_bottom, b->sched_size);
#endif
    emu_ctx_switch(&f.sp, b->sched_sp);
#if EMU_ASAN
    __sanitizer_finish_switch_fiber(f.fake_stack, nullptr, nullptr);
#endif
}

inline void release(Barrier& br) {
    Block* b = tl_block;
    for (int i : br.waiters) b->fib[i].waiting = false;
    br.waiters.clear();
    br.arrived = 0;
}

inline void arrive(Barrier& br) {
    Block* b = tl_block;
    if (++br.arrived >= br.n) {
        release(br);
        return;
    }
    br.waiters.push_back(b->cur);
    b->fib[b->cur].waiting = true;
    switch_to_sched();
}

inline void leave(Barrier& br) {
    --br.n;
    if (br.n > 0 && br.arrived >= br.n) release(br);
}

[[noreturn]] inline void fiber_entry() {
#if EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &tl_block->sched_bottom, &tl_block->sched_size);
#endif
    Block* b = tl_block;
    b->body();
    b = tl_block;
    Fiber& f = b->fib[b->cur];
    leave(b->wave[f.flat / 64]);
    leave(b->bar);
    f.done = true;
#if EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, b->sched_bottom, b->sched_size);
#endif
    void* dummy;
    emu_ctx_switch(&dummy, b->sched_sp);
    __builtin_unreachable();
}

// Runs one block to completion on the calling worker thread.
inline void run_block(Block& b, dim3 bidx, unsigned nt, size_t shmem) {
    tl_block = &b;
    blockIdx = bidx;
    b.fib.resize(nt);
    b.bar = Barrier{(int)nt, 0, {}};
    b.wave.assign((nt + 63) / 64, Barrier{});
    for (unsigned w = 0; w < b.wave.size(); ++w) b.wave[w].n = (int)std::min(64u, nt - 64 * w);
    b.xchg.assign(nt, 0);
    b.lds.assign(shmem, kFill);
    for (unsigned t = 0; t < nt; ++t) {
        Fiber& f = b.fib[t];
        if (f.stack_size != stack_bytes()) {
            f.stack.reset(new unsigned char[stack_bytes()]);
            f.stack_size = stack_bytes();
        }
        f.done = false;
        f.waiting = false;
        f.flat = t;
        f.tid = dim3(t % blockDim.x, (t / blockDim.x) % blockDim.y, t / (blockDim.x * blockDim.y));
        uintptr_t top = ((uintptr_t)(f.stack.get() + f.stack_size)) & ~(uintptr_t)15;
        void** sp = (void**)top;
        *--sp = nullptr;                       // fake return address of fiber_entry
        *--sp = (void*)&fiber_entry;           // where emu_ctx_switch returns to
        for (int r = 0; r < 6; ++r) *--sp = nullptr;
        f.sp = sp;
    }
    unsigned left = nt;
    // EMU_SHUFFLE=<seed>: between two barriers the fibers run in a fresh pseudo-random
    // order on every sweep (not thread 0 first), so a value one thread writes and another
    // reads without a barrier in between shows up as a wrong result.
    static const unsigned long long shuffle_seed = [] {
        const char* e = std::getenv("EMU_SHUFFLE");
        return e ? std::strtoull(e, nullptr, 10) : 0ull;
    }();
    std::vector<unsigned> order(nt);
    for (unsigned t = 0; t < nt; ++t) order[t] = t;
    unsigned long long rs = shuffle_seed * 0x9E3779B97F4A7C15ull + bidx.x * 0xBF58476D1CE4E5B9ull + 1;
    while (left > 0) {
        bool progressed = false;
        if (shuffle_seed) {
            for (unsigned t = nt; t > 1; --t) {
                rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17;
                std::swap(order[t - 1], order[(unsigned)(rs % t)]);
            }
        }
        for (unsigned q = 0; q < nt; ++q) {
            const unsigned t = order[q];
            Fiber& f = b.fib[t];
            if (f.done || f.waiting) continue;
            b.cur = (int)t;
            threadIdx = f.tid;
#if EMU_ASAN
            void* fs = nullptr;
            __sanitizer_start_switch_fiber(&fs, f.stack.get(), f.stack_size);
#endif
            emu_ctx_switch(&b.sched_sp, f.sp);
#if EMU_ASAN
            __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
#endif
            progressed = true;
            if (f.done) --left;
        }
        if (!progressed && left > 0) {
            std::fprintf(stderr, "emu: deadlock in block (%u,%u,%u): %u threads wait at a barrier "
                                 "that the others never reach\n", bidx.x, bidx.y, bidx.z, left);
            std::abort();
        }
    }
    b.cur = -1;
    tl_block = nullptr;
}

template <class T>
T exchange(T v, int src) {
    static_assert(sizeof(T) <= 8, "exchange of at most 8 bytes");
    Block* b = tl_block;
    const unsigned t = b->fib[b->cur].flat;
    const unsigned w = t / 64;
    uint64_t raw = 0;
    std::memcpy(&raw, &v, sizeof(T));
    b->xchg[t] = raw;
    arrive(b->wave[w]);
    const uint64_t got = b->xchg[w * 64 + (unsigned)src];
    arrive(b->wave[w]);
    T r;
    std::memcpy(&r, &got, sizeof(T));
    return r;
}
inline unsigned lane_id() { return tl_block->fib[tl_block->cur].flat % 64; }

// One Block (with its fiber stacks) per worker slot, kept for the whole process.
inline std::vector<std::unique_ptr<Block>>& block_pool(int workers) {
    static std::vector<std::unique_ptr<Block>> blocks;
    static std::mutex m;
    std::lock_guard<std::mutex> g(m);
    while ((int)blocks.size() < workers) blocks.emplace_back(new Block);
    return blocks;
}

template <class F, class... A>
void launch(F kernel, dim3 grid, dim3 block, size_t shmem, A... args) {
    gridDim = grid;
    blockDim = block;
    const unsigned nt = block.x * block.y * block.z;
    const long long nblocks = (long long)grid.x * grid.y * grid.z;
    int workers = 1;
    if (const char* e = std::getenv("EMU_WORKERS")) workers = std::max(1, std::atoi(e));
    workers = (int)std::min<long long>(workers, nblocks);
    std::atomic<long long> next{0};
    std::vector<std::unique_ptr<Block>>& blocks = block_pool(workers);
    auto work_on = [&](int wi) {
        Block& b = *blocks[wi];
        b.body = [&] { kernel(args...); };
        for (long long i; (i = next.fetch_add(1)) < nblocks;) {
            dim3 bi((unsigned)(i % grid.x), (unsigned)((i / grid.x) % grid.y),
                    (unsigned)(i / ((long long)grid.x * grid.y)));
            run_block(b, bi, nt, shmem);
        }
    };
    if (workers == 1) {
        work_on(0);
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < workers; ++w) th.emplace_back(work_on, w);
        for (auto& t : th) t.join();
    }
}
inline void* dyn_lds() { return tl_block->lds.data(); }

}  // namespace emu

inline void __syncthreads() { emu::arrive(emu::tl_block->bar); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }

template <class T> T __shfl_xor(T v, int o, int w = 64) {
    const int l = (int)emu::lane_id();
    return emu::exchange(v, (l & ~(w - 1)) | ((l ^ o) & (w - 1)));
}
template <class T> T __shfl_up(T v, unsigned o, int w = 64) {
    const int l = (int)emu::lane_id(), base = l & ~(w - 1);
    return emu::exchange(v, (l - base) >= (int)o ? l - (int)o : l);
}
template <class T> T __shfl(T v, int s, int w = 64) {
    const int l = (int)emu::lane_id();
    return emu::exchange(v, (l & ~(w - 1)) + (s & (w - 1)));
}
inline int __builtin_amdgcn_readlane(int v, int lane) { return emu::exchange(v, lane); }
// Vote over the lanes of the wavefront that have not returned.
inline unsigned long long __ballot(int pred) {
    emu::Block* b = emu::tl_block;
    const unsigned t = b->fib[b->cur].flat, w = t / 64;
    b->xchg[t] = pred ? 1u : 0u;
    emu::arrive(b->wave[w]);
    unsigned long long m = 0;
    const unsigned n = (unsigned)b->fib.size();
    for (unsigned l = 0; l < 64 && w * 64 + l < n; ++l)
        if (!b->fib[w * 64 + l].done && b->xchg[w * 64 + l]) m |= 1ull << l;
    emu::arrive(b->wave[w]);
    return m;
}
inline unsigned int __popcll(unsigned long long x) { return (unsigned int)__builtin_popcountll(x); }
// Set bits of the mask in the lanes below this one (low and high 32 lanes), plus acc.
inline unsigned int __builtin_amdgcn_mbcnt_lo(unsigned int m, unsigned int acc) {
    const unsigned l = (unsigned)emu::lane_id() & 63u;
    const unsigned long long below = (1ull << l) - 1ull;
    return acc + (unsigned)__builtin_popcount(m & (unsigned)(below & 0xffffffffull));
}
inline unsigned int __builtin_amdgcn_mbcnt_hi(unsigned int m, unsigned int acc) {
    const unsigned l = (unsigned)emu::lane_id() & 63u;
    const unsigned long long below = (1ull << l) - 1ull;
    return acc + (unsigned)__builtin_popcount(m & (unsigned)(below >> 32));
}
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
// All lanes of the wavefront reach this point before any continues (on the GPU the
// lanes run in lockstep; here it orders LDS writes and reads across lanes).
inline void __builtin_amdgcn_wave_barrier() {
    emu::Block* b = emu::tl_block;
    emu::arrive(b->wave[b->fib[b->cur].flat / 64]);
}
inline void __builtin_amdgcn_fence(int, const char*) {}
inline void __builtin_amdgcn_sched_barrier(int) {}   // an instruction-scheduling hint
inline void __builtin_amdgcn_sched_group_barrier(int, int, int) {}
// v_rcp_f64 is an estimate that callers refine with Newton steps; the exact quotient
// is a valid estimate.
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }

inline int __double2loint(double d) { uint64_t u; std::memcpy(&u, &d, 8); return (int)(uint32_t)u; }
inline int __double2hiint(double d) { uint64_t u; std::memcpy(&u, &d, 8); return (int)(uint32_t)(u >> 32); }
inline double __hiloint2double(int hi, int lo) {
    uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    double d; std::memcpy(&d, &u, 8); return d;
}
inline long long __double_as_longlong(double d) { long long u; std::memcpy(&u, &d, 8); return u; }
inline double __longlong_as_double(long long u) { double d; std::memcpy(&d, &u, 8); return d; }

template <class T, class U> T atomicAdd(T* p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
inline double atomicAdd(double* p, double v) {
    uint64_t* q = (uint64_t*)p;
    uint64_t o = __atomic_load_n(q, __ATOMIC_SEQ_CST), n;
    do {
        double od; std::memcpy(&od, &o, 8);
        const double nd = od + v;
        std::memcpy(&n, &nd, 8);
    } while (!__atomic_compare_exchange_n(q, &o, n, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
    double od; std::memcpy(&od, &o, 8);
    return od;
}
inline float atomicAdd(float* p, float v) {
    uint32_t* q = (uint32_t*)p;
    uint32_t o = __atomic_load_n(q, __ATOMIC_SEQ_CST), n;
    do {
        float of; std::memcpy(&of, &o, 4);
        const float nf = of + v;
        std::memcpy(&n, &nf, 4);
    } while (!__atomic_compare_exchange_n(q, &o, n, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
    float of; std::memcpy(&of, &o, 4);
    return of;
}
template <class T, class U> T atomicOr(T* p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T> T atomicMax(T* p, T v) {
    T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (o < v && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return o;
}
template <class T> T atomicMin(T* p, T v) {
    T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (o > v && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return o;
}
template <class T> T atomicExch(T* p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T> T atomicCAS(T* p, T c, T v) {
    __atomic_compare_exchange_n(p, &c, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return c;
}

// ---- runtime ----
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorLaunchFailure = 719 };
typedef struct emu_stream* hipStream_t;
typedef struct emu_event* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
enum { hipStreamNonBlocking = 1 };

inline const char* hipGetErrorString(hipError_t) { return "emulated HIP error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
template <class T> hipError_t hipMalloc(T** p, size_t n) {
    void* q = std::malloc(n ? n : 1);
    if (!q) return hipErrorInvalidValue;
    std::memset(q, emu::kFill, n);
    *p = (T*)q;
    return hipSuccess;
}
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
template <class T> hipError_t hipHostMalloc(T** p, size_t n, unsigned = 0) { return hipMalloc(p, n); }
inline hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
// device symbols are host objects here: their address is the object's
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t = nullptr) { std::memset(p, v, n); return hipSuccess; }
inline hipError_t hipMemset(void* p, int v, size_t n) { std::memset(p, v, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t = nullptr) {
    std::memmove(d, s, n); return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    static int one_stream;   // every launch runs synchronously; the handle is only a token
    *s = (hipStream_t)&one_stream;
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = nullptr; return hipSuccess; }
enum { hipEventDefault = 0, hipEventDisableTiming = 2 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = nullptr; return hipSuccess; }
// kernels run to completion at launch, so every stream is always idle
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
// HIP graphs are not emulated: a capture fails, so callers keep to eager launches
typedef struct emu_graph* hipGraph_t;
typedef struct emu_graph_exec* hipGraphExec_t;
enum { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
       hipStreamCaptureModeRelaxed = 2 };
typedef int hipStreamCaptureMode;
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorInvalidValue; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = nullptr; return hipErrorInvalidValue; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, void*, void*, size_t) { *e = nullptr; return hipErrorInvalidValue; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
inline hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
// an emulated device of 8 compute units holding 2 workgroups each (EMU_CUS overrides 8)
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
    const char* e = std::getenv("EMU_CUS");
    *v = e ? std::atoi(e) : 8;
    return hipSuccess;
}
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) {
    *n = 2;
    return hipSuccess;
}

#define HIP_DYNAMIC_SHARED(type, var) type* var = (type*)emu::dyn_lds();
#define hipLaunchKernelGGL(K, G, B, S, ST, ...) ((void)(ST), emu::launch(K, dim3(G), dim3(B), (size_t)(S), __VA_ARGS__))
