#include "profile.hpp"

#include "tree.hpp"

namespace spand {

static const char* const kFamilyNames[KF_COUNT] = {
    "fill", "scatter", "copy", "potrf", "potrf_panel", "potrf_trail", "trsm", "schur", "rrqr", "rrqr_init", "rrqr_step", "rrqr_wave",
    "rrqr_finish", "larft", "perm", "trmv", "gemv", "qapply", "merge", "spmv", "krylov_vec", "getrf", "rrqr_reg",
};

const char* kernel_family_name(int f) { return (f >= 0 && f < KF_COUNT) ? kFamilyNames[f] : nullptr; }

KernelProfiler::~KernelProfiler() {
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : free_) (void)hipEventDestroy(e);
}

hipEvent_t KernelProfiler::get_event() {
    if (!free_.empty()) {
        hipEvent_t e = free_.back();
        free_.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

int KernelProfiler::start(int fam, hipStream_t st) {
    if (!on(fam)) return -1;
    Pending p{fam, get_event(), get_event(), 0.0, 0.0, 1};
    hip_check(hipEventRecord(p.a, st), "hipEventRecord");
    pending_.push_back(p);
    return (int)pending_.size() - 1;
}

void KernelProfiler::stop(int slot, hipStream_t st, double bytes, double flops, int launches) {
    if (slot < 0) return;
    Pending& p = pending_[slot];
    p.bytes = bytes;
    p.flops = flops;
    p.launches = launches;
    hip_check(hipEventRecord(p.b, st), "hipEventRecord");
}

void KernelProfiler::set_work(int slot, double bytes, double flops) {
    if (slot < 0 || slot >= (int)pending_.size()) return;
    pending_[slot].bytes = bytes;
    pending_[slot].flops = flops;
}

void KernelProfiler::collect() {
    for (auto& p : pending_) {
        hip_check(hipEventSynchronize(p.b), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        KernelTotals& t = acc_[p.fam];
        t.launches += p.launches;
        t.ms += ms;
        t.bytes += p.bytes;
        t.flops += p.flops;
        free_.push_back(p.a);
        free_.push_back(p.b);
    }
    pending_.clear();
}

void KernelProfiler::reset() {
    collect();
    for (auto& a : acc_) a = KernelTotals{};
}

}  // namespace spand
