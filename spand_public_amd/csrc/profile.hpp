// Optional per-launch kernel timing with HIP events recorded on the launch stream, plus the
// algorithmic bytes and flops of every timed launch (the numerators of a roofline figure).
// Off by default; when on, every library launch is bracketed by two events and the totals
// are accumulated per kernel family at the next collect() (which waits for the events).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace spand {

enum KernelFamily {
    KF_FILL = 0,
    KF_SCATTER,
    KF_COPY,
    KF_POTRF,
    KF_POTRF_PANEL,
    KF_POTRF_TRAIL,
    KF_TRSM,
    KF_SCHUR,
    KF_RRQR,
    KF_RRQR_INIT,
    KF_RRQR_STEP,
    KF_RRQR_WAVE,
    KF_RRQR_FINISH,
    KF_LARFT,
    KF_PERM,
    KF_TRMV,
    KF_GEMV,
    KF_QAPPLY,
    KF_MERGE,
    KF_SPMV,
    KF_KRYLOV_VEC,
    KF_GETRF,
    KF_RRQR_REG,
    KF_COUNT
};

const char* kernel_family_name(int f);

struct KernelTotals {
    long long launches = 0;
    double ms = 0, bytes = 0, flops = 0;
};

class KernelProfiler {
  public:
    ~KernelProfiler();
    // bit f of mask enables family f (0: off)
    void set_mask(unsigned long long mask) { mask_ = mask; }
    bool on() const { return mask_ != 0; }
    bool on(int fam) const { return (mask_ >> fam) & 1ull; }
    // start() before a launch, stop() after it; both return/take a slot (-1 when the family
    // is off). A slot may bracket `launches` consecutive launches of its family (a batch);
    // its time then covers the batch, gaps between the launches included.
    int start(int fam, hipStream_t st);
    void stop(int slot, hipStream_t st, double bytes, double flops, int launches = 1);
    // Work of a launch known only after it ran (sizes decided on the device).
    void set_work(int slot, double bytes, double flops);
    void collect();  // waits for pending events, accumulates, recycles the events
    void reset();
    KernelTotals totals(int fam) const { return (fam >= 0 && fam < KF_COUNT) ? acc_[fam] : KernelTotals{}; }

  private:
    struct Pending {
        int fam;
        hipEvent_t a, b;
        double bytes, flops;
        int launches;
    };
    hipEvent_t get_event();
    std::vector<hipEvent_t> free_;
    std::vector<Pending> pending_;
    KernelTotals acc_[KF_COUNT];
    unsigned long long mask_ = 0;
};

// Brackets one launcher call: `hipError_t` of the launcher is returned unchanged.
template <typename F>
hipError_t timed(KernelProfiler& p, int fam, hipStream_t st, double bytes, double flops, F&& launch,
                 int* slot_out = nullptr) {
    const int s = p.start(fam, st);
    hipError_t e = launch();
    p.stop(s, st, bytes, flops);
    if (slot_out) *slot_out = s;
    return e;
}

}  // namespace spand
