// C ABI of the coherent-search engine: context, buffers, batching and the
// entry points declared in include/coherent_search.h.
#include "coherent_search.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <limits>
#include <cstring>
#include <future>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "cs_fused.hpp"
#include "cs_geo.hpp"
#include "cs_mfma.hpp"
#include "cs_gpu.hpp"
#include "cs_plan.hpp"
#include "cs_stats.hpp"
#include "cs_triage.hpp"

using namespace cs;

namespace {

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

constexpr int kBankW[12] = {1, 2, 3, 4, 6, 9, 13, 19, 28, 42, 63, 94};
constexpr int64_t kBatchTrials = (int64_t)1 << 21;   // plain path: trials per batch (HBM scratch bound)
constexpr int64_t kBatchTrialsFused = (int64_t)1 << 25;   // fused path: ~1 GB of bounds + samples per batch
// split-BF16 path: blocks per batch (6 folds: 3.5 GB of sigma samples and
// 0.8 GB of bounds at 2048-trial blocks), one batch at 2^25 points
constexpr int64_t kBatchBlocksMfma = 16384;
constexpr int kMaxHarm = 64;                         // device path: profiles of <= 128 bins

}  // namespace

struct cs_ctx {
    // parameters
    int32_t nharms = 0, m = 0, numbetween = 16, precision = 0;
    double hidr = 0.5, threshold = 8.0, fsp = 1.5, maxfrac = 0.3, medmargin = 2.0;
    std::vector<int> decims;
    int64_t Nprof = 2048;
    int device = 0;
    int mode = CS_MODE_PLAIN;
    hipStream_t st = nullptr;
    std::string err;
    // spectrum
    DBuf amps;
    int64_t namps = 0, N = 0;
    double T = 0.0;
    // tables
    DBuf harm, row, carry, W32, A32, W64, A64;
    std::map<int, DBuf> twid;
    // scratch
    DBuf ftp, prof, nvalid, sigma, surv, counters, cands, widths, rfund, rowok, mout, mraw, samples;
    DBuf rflags;    // (block, fold) pairs of a batch whose sigma is refined in FP64, + count
    DBuf rlist;     // the flagged pairs, each once
    DBuf rsamp;     // FP64 samples of the first kRefineCap listed pairs
    DBuf probe_buf; // cs_probe_split_bf16: the copied-out stacked amplitudes
    DBuf bandtab;   // block_metrics(:fft): per-harmonic lobin, then fine-grid length
    DBuf cand_r, cand_h, cand_w, cand_prof;   // candidate profiles: inputs, widths, outputs
    // metric statistics: every metric of the last statistics search
    // ([fold][trial]), histograms, block records, windows, normalisation tables
    bool stats_on = false;
    DBuf st_metric, st_cnt, st_blk, st_win, st_norm;
    struct {
        bool valid = false;
        int nf = 0, nwin = 0, nb = 0;
        double lo = 0, hi = 0, r_lo = 0, lodr = 0, T = 0;
        int64_t total = 0, Nprof = 0, nchunks = 0, N = 0;
        std::vector<int> ks, Hks;            // folds in search order
        std::vector<double> edges;           // [fold][nwin + 1]
        std::vector<int32_t> win, nv;        // [fold][block], [block][fold]
        std::vector<BlockStatDev> blk;       // [fold][block]
        std::vector<uint32_t> cnt;           // [fold][window][bin]
    } sh;
    DBuf pyr, tfeat, mlp_w, mlp_x, mlp_p;   // block triage: power pyramid, features, MLP buffers
    DBuf m_whi, m_wlo, m_gA, m_hdesc, m_chi, m_clo, m_desc, m_rowmask, m_scol, m_nv16;
    std::vector<uint8_t> m_desc_last;   // host copy of what m_desc holds
    // split-BF16 batches: identity list (u32 i at i, m_ident_n entries), and
    // for a block subset its block list, slot table and per-batch tile lists
    DBuf m_ident, m_sel, m_slot, m_tiles;
    int64_t m_ident_n = 0;
    DBuf spl;                  // pre-split spectrum planes of the MFMA mode
    bool spl_valid = false;    // spl matches the current spectrum
    int ncu = 0;               // compute units of the device
    int64_t tune_batch = 0;    // CS_TUNE_BATCH_BLOCKS (0: default)
    // multi-file search: spectrum view (a file's resident copy instead of
    // `amps`), two pinned staging buffers and the upload stream
    const void* amps_view = nullptr;
    DBuf sweep_amps, sweep_cands, sweep_ctr;   // cs_search_files: resident spectra, results
    std::vector<hipEvent_t> sweep_up;          // per file: its upload has completed
    hipEvent_t sweep_ev[2] = {nullptr, nullptr};
    void* pin[2] = {nullptr, nullptr};
    size_t pin_bytes = 0;
    void* hout = nullptr;   // pinned: a search's counters, then its first kOutRecords candidates
    hipStream_t cst = nullptr;
    int tune_wg = 0;           // CS_TUNE_WORKGROUPS (0: one per CU)
    int tune_order = 0;        // CS_TUNE_TILE_ORDER (0: contiguous tile ranges per XCD)
    int tune_generic = 0;      // CS_TUNE_GENERIC (1: no fold-set kernel)
    int tune_noprefetch = 0;   // CS_TUNE_NO_PREFETCH (1: no LDS prefetch of conj(A))
    bool use_graph = false;    // CS_TUNE_GRAPHS: replay a captured launch sequence
    std::string tune_cache;    // cs_set_tuning_cache: cs_autotune's cache for cs_search_files
    hipGraph_t g_graph = nullptr;
    hipGraphExec_t g_exec = nullptr;
    std::vector<uint8_t> g_key;   // every launch argument the graph was captured with
    int64_t g_captures = 0;
    // host-built tables are kept while the grid they were built for is unchanged
    bool tab_valid = false;    // plans on the device for (tab_rlo, tab_f64)
    double tab_rlo = 0.0;
    bool tab_f64 = false;
    TabDev tab{};
    bool mfma_valid = false;   // MFMA operand tables for mfma_rlo and mfma_nbs (lengths, then widest widths)
    double mfma_rlo = 0.0;
    std::vector<int> mfma_nbs;
    MfmaDev mfma{};
    bool nv_valid = false;     // per-block valid counts for (nv_rlo, nv_total, nv_N)
    double nv_rlo = 0.0;
    int64_t nv_total = 0, nv_N = 0;
    std::vector<hipEvent_t> ev;
    double timings[8] = {0};
    // block sigmas of the last search: [block][fold] in `sigma`, for a search
    // over every block of its range
    int64_t sig_blocks = 0;
    int sig_nf = 0;
    bool sig_sparse = false;
    // this device's copy of the allocation-free entry point's scratch array
    // (its address from hipGetSymbolAddress at context creation)
    void* scratch_dev = nullptr;
    hipError_t scratch_err = hipSuccess;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
};

#define HIPCHK(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return (ctx)->fail(e_ == hipErrorOutOfMemory ? CS_ENOMEM : CS_EDEVICE,          \
                               std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)

namespace {

int ensure(cs_ctx* c, DBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return CS_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return c->fail(CS_EDEVICE, std::string("hipFree: ") + hipGetErrorString(e));
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return c->fail(CS_ENOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return CS_OK;
}

void release(DBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

template <typename T>
int upload(cs_ctx* c, DBuf& b, const std::vector<T>& v) {
    int s = ensure(c, b, v.size() * sizeof(T));
    if (s) return s;
    if (!v.empty())
        HIPCHK(c, hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

// Upload the interpolation tables for a grid anchored at r_lo and fill `tb`.
int set_tables(cs_ctx* c, double r_lo, bool weights_f64, TabDev& tb, int64_t last_trial) {
    if (last_trial >= ((int64_t)1 << 32))
        return c->fail(CS_EINVAL, "more than 2^32 trials in one search");
    if (c->tab_valid && c->tab_rlo == r_lo && c->tab_f64 == weights_f64) {
        tb = c->tab;
        return CS_OK;
    }
    c->tab_valid = false;
    Plans pl = build_plans(c->nharms, c->m, c->hidr, r_lo);
    if (pl.rational) {
        for (const HarmDev& d : pl.harm)
            if ((uint64_t)d.q * (uint64_t)(d.hp > 0 ? d.hp : 1) >= ((uint64_t)1 << 32))
                return c->fail(CS_EINVAL, "trial grid q*h*pnum exceeds 2^32");
    }
    if (last_trial >= ((int64_t)1 << 32))
        return c->fail(CS_EINVAL, "more than 2^32 trials in one search");
    int s;
    if ((s = upload(c, c->harm, pl.harm))) return s;
    if ((s = upload(c, c->row, pl.row))) return s;
    if ((s = upload(c, c->carry, pl.carry))) return s;
    if ((s = upload(c, c->W32, pl.W32))) return s;
    if ((s = upload(c, c->A32, pl.A32))) return s;
    if ((s = upload(c, c->W64, pl.W64))) return s;
    if ((s = upload(c, c->A64, pl.A64))) return s;
    tb.harm = (const HarmDev*)c->harm.p;
    tb.row = (const int32_t*)c->row.p;
    tb.carry = (const uint8_t*)c->carry.p;
    tb.W32 = (const float*)c->W32.p;
    tb.A32 = (const float2*)c->A32.p;
    tb.W64 = (const double*)c->W64.p;
    tb.A64 = (const double2*)c->A64.p;
    tb.lodr = pl.lodr;
    tb.m = c->m;
    tb.nharms = c->nharms;
    tb.weights_f64 = weights_f64 ? 1 : 0;
    tb.pad = 0;
    c->tab = tb;
    c->tab_rlo = r_lo;
    c->tab_f64 = weights_f64;
    c->tab_valid = true;
    return CS_OK;
}

int get_twid(cs_ctx* c, int nbins, const double2** out) {
    auto it = c->twid.find(nbins);
    if (it == c->twid.end()) {
        std::vector<double> v(2 * nbins);
        for (int e = 0; e < nbins; ++e) {
            double a = 2.0 * (double)e / (double)nbins;
            v[2 * e] = cospi_d(a);
            v[2 * e + 1] = sinpi_d(a);
        }
        DBuf b;
        int s = upload(c, b, v);
        if (s) return s;
        it = c->twid.emplace(nbins, b).first;
    }
    *out = (const double2*)it->second.p;
    return CS_OK;
}

SpecDev spec_of(cs_ctx* c) {
    SpecDev sp;
    sp.amps = (const float2*)(c->amps_view ? c->amps_view : c->amps.p);
    sp.namps = c->namps;
    sp.N = c->N;
    sp.Nhalf = c->N / 2;
    return sp;
}

std::vector<int> ladder(cs_ctx* c, int nbins, int k) {
    return ladder_boxcar_widths(nbins, k, c->decims, c->fsp, c->maxfrac);
}

uint32_t bank_mask(const std::vector<int>& ws) {
    uint32_t mask = 0;
    for (int w : ws) {
        int a = -1;
        for (int i = 0; i < 12; ++i) if (kBankW[i] == w) a = i;
        if (a < 0) return 0;
        mask |= 1u << a;
    }
    return mask;
}

// nvalid: the prefix j of block trials with k*(rstart + j*lodr) < N/2
// (decim_pass!, search.jl:1679-1691); the predicate is monotone in j.
int64_t nvalid_of(int k, double rstart, double lodr, int64_t n, double nyq) {
    if (k == 1) return n;
    auto ok = [&](int64_t j) { return (double)k * (rstart + (double)j * lodr) < nyq; };
    if (n == 0 || ok(n - 1)) return n;
    int64_t lo = 0, hi = n - 1;   // ok(hi) false
    if (!ok(0)) return 0;
    while (hi - lo > 1) {
        int64_t mid = lo + (hi - lo) / 2;
        if (ok(mid)) lo = mid;
        else hi = mid;
    }
    return hi;
}

int check_spectrum(cs_ctx* c) {
    if ((!c->amps.p && !c->amps_view) || c->namps <= 0)
        return c->fail(CS_EINVAL, "no spectrum set (cs_set_spectrum)");
    return CS_OK;
}

}  // namespace

// ---- scratch of cs_snr_metrics_noalloc ----------------------------------------
// One file-scope device array: every GPU the library runs on holds its own
// copy until the process ends, at an address of its own, looked up once per
// device (hipGetSymbolAddress with that device current) and kept per device.
// Host code never passes `scratch` itself to a kernel: on the host that name
// is a stand-in, not a device address.  Layout: [0, 8) the sigma, [256, 512)
// the width bank (int32), [kScratchHead, ...) up to kScratchSamples sigma
// samples (double).
constexpr int64_t kScratchSamples = (int64_t)1 << 20;
constexpr size_t kScratchHead = 4096;
constexpr size_t kScratchBytes = kScratchHead + (size_t)kScratchSamples * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_at[kScratchDevices] = {};

// this device's address of `scratch` (the caller has made `device` current)
hipError_t scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_at[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_at[device] = ptr;
    }
    *out = g_scratch_at[device];
    return hipSuccess;
}

// a boxcar width bank passed by value (no host-to-device copy)
struct WidthBank {
    int32_t n;
    int32_t w[64];
};

__global__ void k_put_widths(WidthBank wb, int32_t* dst) {
    if ((int)threadIdx.x < wb.n) dst[threadIdx.x] = wb.w[threadIdx.x];
}

}  // namespace

extern "C" {

const char* cs_version(void) { return "coherentsearch-mi355x 0.1 (gfx950)"; }

const char* cs_last_error(const cs_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cs_ctx_create(const cs_params* p, int32_t blocksize, int32_t device, cs_ctx** out) {
    if (!out || !p) return CS_EINVAL;
    *out = nullptr;
    if (p->m % 2 != 0 || p->m <= 0) return CS_EINVAL;
    if (p->nharms < 1 || p->nharms > kMaxHarm) return CS_EINVAL;
    if (p->precision != 0 && p->precision != 1) return CS_EINVAL;
    if (p->ndecim < 0 || p->ndecim > kMaxFolds - 1 + 1) return CS_EINVAL;
    cs_ctx* c = new cs_ctx();
    c->nharms = p->nharms;
    c->m = p->m;
    c->numbetween = p->numbetween;
    c->precision = p->precision;
    c->hidr = p->hidr;
    c->threshold = p->threshold;
    c->fsp = p->boxcar_fsp;
    c->maxfrac = p->boxcar_maxfrac;
    c->medmargin = p->boxcar_medmargin;
    for (int i = 0; i < p->ndecim; ++i) c->decims.push_back(p->decimations[i]);
    if (c->decims.empty()) c->decims.push_back(1);
    int nf = 1;
    for (int k : c->decims) {
        if (k < 1 || c->nharms / k < 1) { delete c; return CS_EINVAL; }
        if (k > 1) ++nf;
    }
    if (nf > kMaxFolds) { delete c; return CS_EINVAL; }
    c->Nprof = std::max<int64_t>(1, blocksize);
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return CS_EDEVICE;
    }
    // the allocation-free entry point's scratch on this device; a failure is
    // reported by that entry point only
    c->scratch_err = scratch_address(device, &c->scratch_dev);
    *out = c;
    return CS_OK;
}

int cs_ctx_destroy(cs_ctx* c) {
    if (!c) return CS_OK;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    DBuf* bufs[] = {&c->amps, &c->harm, &c->row, &c->carry, &c->W32, &c->A32, &c->W64, &c->A64,
                    &c->ftp, &c->prof, &c->nvalid, &c->sigma, &c->surv, &c->counters, &c->cands,
                    &c->widths, &c->rfund, &c->rowok, &c->mout, &c->mraw, &c->samples, &c->bandtab,
                    &c->cand_r, &c->cand_h, &c->cand_w, &c->cand_prof,
                    &c->st_metric, &c->st_cnt, &c->st_blk, &c->st_win, &c->st_norm,
                    &c->pyr, &c->tfeat, &c->mlp_w, &c->mlp_x, &c->mlp_p,
                    &c->m_whi, &c->m_wlo, &c->m_gA, &c->m_hdesc, &c->m_chi, &c->m_clo,
                    &c->m_desc, &c->m_rowmask, &c->m_scol, &c->m_nv16, &c->m_ident, &c->m_sel,
                    &c->m_slot,
                    &c->m_tiles, &c->spl, &c->rflags, &c->rlist, &c->rsamp, &c->probe_buf, &c->sweep_amps,
                    &c->sweep_cands, &c->sweep_ctr};
    for (DBuf* b : bufs) release(*b);
    for (void* p : c->pin)
        if (p) (void)hipHostFree(p);
    if (c->hout) (void)hipHostFree(c->hout);
    for (hipEvent_t e : c->sweep_up)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->sweep_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->cst) (void)hipStreamDestroy(c->cst);
    for (auto& kv : c->twid) release(kv.second);
    for (hipEvent_t ev : c->ev) (void)hipEventDestroy(ev);
    if (c->g_exec) (void)hipGraphExecDestroy(c->g_exec);
    if (c->g_graph) (void)hipGraphDestroy(c->g_graph);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return CS_OK;
}

int cs_set_mode(cs_ctx* c, int32_t mode) {
    if (!c) return CS_EINVAL;
    if (mode != CS_MODE_PLAIN && mode != CS_MODE_FUSED && mode != CS_MODE_FUSED_VALU)
        return c->fail(CS_EINVAL, "unknown mode");
    c->mode = mode;
    return CS_OK;
}

int cs_set_tuning(cs_ctx* c, int32_t knob, int64_t value) {
    if (!c) return CS_EINVAL;
    if (value < 0) return c->fail(CS_EINVAL, "tuning value must be >= 0");
    if (knob == CS_TUNE_BATCH_BLOCKS) {
        c->tune_batch = value;
    } else if (knob == CS_TUNE_WORKGROUPS) {
        if (value > 65536) return c->fail(CS_EINVAL, "at most 65536 workgroups");
        c->tune_wg = (int)value;
    } else if (knob == CS_TUNE_GRAPHS) {
        if (value > 1) return c->fail(CS_EINVAL, "CS_TUNE_GRAPHS takes 0 or 1");
        c->use_graph = value == 1;
    } else if (knob == CS_TUNE_TILE_ORDER) {
        if (value > 1) return c->fail(CS_EINVAL, "CS_TUNE_TILE_ORDER takes 0 or 1");
        c->tune_order = (int)value;
    } else if (knob == CS_TUNE_GENERIC) {
        if (value > 1) return c->fail(CS_EINVAL, "CS_TUNE_GENERIC takes 0 or 1");
        c->tune_generic = (int)value;
    } else if (knob == CS_TUNE_NO_PREFETCH) {
        if (value > 1) return c->fail(CS_EINVAL, "CS_TUNE_NO_PREFETCH takes 0 or 1");
        c->tune_noprefetch = (int)value;
    } else {
        return c->fail(CS_EINVAL, "unknown tuning knob");
    }
    return CS_OK;
}

// cs_autotune: the cache line of a search geometry (parameters, trial range,
// spectrum length, device architecture)
int cs_get_tuning(const cs_ctx* c, int32_t knob, int64_t* value) {
    if (!c || !value) return CS_EINVAL;
    if (knob == CS_TUNE_BATCH_BLOCKS) *value = c->tune_batch;
    else if (knob == CS_TUNE_WORKGROUPS) *value = c->tune_wg;
    else if (knob == CS_TUNE_GRAPHS) *value = c->use_graph ? 1 : 0;
    else if (knob == CS_TUNE_TILE_ORDER) *value = c->tune_order;
    else if (knob == CS_TUNE_GENERIC) *value = c->tune_generic;
    else if (knob == CS_TUNE_NO_PREFETCH) *value = c->tune_noprefetch;
    else return CS_EINVAL;
    return CS_OK;
}

static std::string tune_key(cs_ctx* c, double r_lo, double r_hi) {
    hipDeviceProp_t prop;
    const char* arch = hipGetDeviceProperties(&prop, c->device) == hipSuccess ? prop.gcnArchName : "?";
    std::string k = "nh=" + std::to_string(c->nharms) + " m=" + std::to_string(c->m) + " dec=";
    for (int d : c->decims) k += std::to_string(d) + ",";
    char buf[256];
    std::snprintf(buf, sizeof(buf), " hidr=%.17g nprof=%lld mode=%d namps=%lld rlo=%.17g rhi=%.17g arch=",
                  c->hidr, (long long)c->Nprof, c->mode, (long long)c->namps, r_lo, r_hi);
    k += buf;
    for (const char* q = arch; *q && *q != ':'; ++q) k += (*q == ' ' || *q == '\t') ? '_' : *q;
    return k;
}

// The launch shape of a cache line "key<TAB>batch workgroups ms order" whose
// key is `key` (false for another key or a malformed entry).
static bool tune_parse(const std::string& l, const std::string& key, int64_t& b, int& w, int& o) {
    const size_t tab = l.find('\t');
    if (tab == std::string::npos || l.compare(0, tab, key) != 0 || tab != key.size()) return false;
    long long bb = 0, ww = 0, oo = 0;
    double msv = 0.0;
    if (std::sscanf(l.c_str() + tab + 1, "%lld %lld %lf %lld", &bb, &ww, &msv, &oo) < 2 || bb < 0 ||
        ww < 0 || ww > 65536 || (oo != 0 && oo != 1))
        return false;
    b = bb;
    w = (int)ww;
    o = (int)oo;
    return true;
}

// The cached launch shape of the context's geometry over band [r_lo, r_hi]
// in the cache file at `path`, if it holds one.
static bool tune_lookup(cs_ctx* c, const std::string& path, double r_lo, double r_hi, int64_t& b,
                        int& w, int& o) {
    FILE* fh = std::fopen(path.c_str(), "r");
    if (!fh) return false;
    const std::string key = tune_key(c, r_lo, r_hi);
    char ln[1024];
    bool found = false;
    while (!found && std::fgets(ln, sizeof(ln), fh)) {
        std::string l(ln);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        found = tune_parse(l, key, b, w, o);
    }
    std::fclose(fh);
    return found;
}

static int tune_time(cs_ctx* c, double r_lo, double r_hi, double thr, int reps, double* ms) {
    std::vector<cs_cand> buf(1 << 12);
    std::vector<double> t;
    for (int r = 0; r <= reps; ++r) {   // the first search warms the caches
        int64_t n = 0;
        int st = cs_search(c, r_lo, r_hi, thr, buf.data(), (int64_t)buf.size(), &n);
        if (st == CS_ENOSPC) {
            buf.resize((size_t)n);
            st = cs_search(c, r_lo, r_hi, thr, buf.data(), (int64_t)buf.size(), &n);
        }
        if (st != CS_OK) return st;
        if (r > 0) t.push_back(c->timings[6]);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return CS_OK;
}

int cs_set_tuning_cache(cs_ctx* c, const char* path) {
    if (!c) return CS_EINVAL;
    c->tune_cache = path ? path : "";
    return CS_OK;
}

int cs_autotune(cs_ctx* c, double r_lo, double r_hi, double threshold, const char* cache_path,
                int32_t reps, int64_t* out_batch, int32_t* out_wg) {
    if (!c) return CS_EINVAL;
    int s = check_spectrum(c);
    if (s) return s;
    if (reps < 1 || reps > 50) return c->fail(CS_EINVAL, "reps must be in 1..50");
    auto report = [&]() {
        if (out_batch) *out_batch = c->tune_batch;
        if (out_wg) *out_wg = c->tune_wg;
        return CS_OK;
    };
    const std::string key = tune_key(c, r_lo, r_hi);
    std::vector<std::string> lines;
    if (cache_path && *cache_path) {
        if (FILE* fh = std::fopen(cache_path, "r")) {
            char ln[1024];
            while (std::fgets(ln, sizeof(ln), fh)) {
                std::string l(ln);
                while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
                if (l.empty()) continue;
                const size_t tab = l.find('\t');
                if (tab != std::string::npos && l.compare(0, tab, key) == 0 && tab == key.size()) {
                    int64_t b = 0;
                    int w = 0, o = 0;
                    if (tune_parse(l, key, b, w, o)) {
                        std::fclose(fh);
                        c->tune_batch = b;
                        c->tune_wg = w;
                        c->tune_order = o;
                        return report();
                    }
                    continue;   // malformed entry: measured again and replaced
                }
                lines.push_back(l);
            }
            std::fclose(fh);
        }
    }
    // candidates.  Split-BF16 mode: persistent grid sizes around one
    // workgroup per CU, batch sizes dividing the search into 1, 2 or 4
    // batches, and the tile order (contiguous ranges per XCD or round-robin).
    // Plain and VALU-fused modes: blocks per batch (the default, half and
    // twice).  Every candidate returns the same results (the knobs only
    // schedule the trials).
    const double lodr = c->hidr / c->nharms;
    const double span = std::floor((r_hi - r_lo) / lodr);
    const int64_t total = span >= 0 ? (int64_t)span + 1 : 0;
    const int64_t nchunks = (total + c->Nprof - 1) / c->Nprof;
    const int ncu = c->ncu > 0 ? c->ncu : 256;
    std::vector<int> wgs{0}, orders{0};
    std::vector<int64_t> bats{0};
    if (c->mode == CS_MODE_FUSED) {
        wgs = {0, std::max(8, ncu * 3 / 4 / 8 * 8), std::max(8, ncu / 2 / 8 * 8)};
        orders = {0, 1};
        for (int64_t d : {2, 4})
            if (nchunks >= d) bats.push_back((nchunks + d - 1) / d);
    } else {
        const int64_t def = std::max<int64_t>(
            1, (c->mode == CS_MODE_FUSED_VALU ? kBatchTrialsFused : kBatchTrials) / c->Nprof);
        for (int64_t bt : {def / 2, def * 2})
            if (bt >= 1 && bt < nchunks * 2) bats.push_back(bt);
    }
    double best = INFINITY;
    int64_t bb = 0;
    int bw = 0, bo = 0;
    for (int o : orders)
        for (int w : wgs)
            for (int64_t bt : bats) {
                c->tune_wg = w;
                c->tune_batch = bt;
                c->tune_order = o;
                double ms = 0.0;
                if ((s = tune_time(c, r_lo, r_hi, threshold, reps, &ms))) {
                    c->tune_wg = 0;
                    c->tune_batch = 0;
                    c->tune_order = 0;
                    return s;
                }
                if (ms < best) {
                    best = ms;
                    bb = bt;
                    bw = w;
                    bo = o;
                }
            }
    c->tune_batch = bb;
    c->tune_wg = bw;
    c->tune_order = bo;
    if (cache_path && *cache_path) {   // temp file + rename, as the FFTW wisdom export
        char val[128];
        std::snprintf(val, sizeof(val), "\t%lld %d %.4f %d", (long long)bb, bw, best, bo);
        lines.push_back(key + val);
        const std::string tmp = std::string(cache_path) + "." + std::to_string((long long)getpid()) + ".tmp";
        bool ok = false;
        if (FILE* fh = std::fopen(tmp.c_str(), "w")) {
            ok = true;
            for (const std::string& l : lines) ok = ok && std::fprintf(fh, "%s\n", l.c_str()) > 0;
            ok = (std::fclose(fh) == 0) && ok;
        }
        if (!ok || std::rename(tmp.c_str(), cache_path) != 0) {
            std::remove(tmp.c_str());
            return c->fail(CS_EINVAL, std::string("cannot write the tuning cache ") + cache_path);
        }
    }
    return report();
}

int cs_last_timings(const cs_ctx* c, double* ms8) {
    if (!c || !ms8) return CS_EINVAL;
    for (int i = 0; i < 8; ++i) ms8[i] = c->timings[i];
    return CS_OK;
}

int cs_last_sigmas(cs_ctx* c, double* out, int64_t cap, int64_t* n_out) {
    if (!c || !n_out) return CS_EINVAL;
    *n_out = 0;
    if (c->sig_blocks <= 0) return c->fail(CS_EINVAL, "no completed search on this context");
    if (c->sig_sparse)
        return c->fail(CS_EINVAL, "the last search covered selected blocks only");
    const int64_t n = c->sig_blocks * c->sig_nf;
    *n_out = n;
    if (!out) return CS_OK;
    if (cap < n) return c->fail(CS_ENOSPC, "sigma buffer too small");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(out, c->sigma.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return CS_OK;
}

static int set_spectrum_impl(cs_ctx* c, const void* src, hipMemcpyKind kind, int64_t namps,
                             int64_t N, double T) {
    if (!c) return CS_EINVAL;
    if (namps <= 0 || N <= 0 || !(T > 0)) return c->fail(CS_EINVAL, "invalid spectrum dimensions");
    HIPCHK(c, hipSetDevice(c->device));
    int s = ensure(c, c->amps, (size_t)namps * sizeof(float2));
    if (s) return s;
    HIPCHK(c, hipMemcpyAsync(c->amps.p, src, (size_t)namps * sizeof(float2), kind, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->amps_view = nullptr;
    c->namps = namps;
    c->N = N;
    c->T = T;
    c->spl_valid = false;
    return CS_OK;
}

int cs_set_spectrum(cs_ctx* c, const float* amps, int64_t namps, int64_t N, double T) {
    return set_spectrum_impl(c, amps, hipMemcpyHostToDevice, namps, N, T);
}

int cs_set_spectrum_device(cs_ctx* c, const void* d_amps, int64_t namps, int64_t N, double T) {
    return set_spectrum_impl(c, d_amps, hipMemcpyDeviceToDevice, namps, N, T);
}

int cs_interp_rows(cs_ctx* c, double r_lo, int64_t t0, int64_t n, int32_t weights_f64,
                   double* out) {
    if (!c || !out) return CS_EINVAL;
    if (n < 0 || t0 < 0) return c->fail(CS_EINVAL, "n and t0 must be non-negative");
    int s = check_spectrum(c);
    if (s) return s;
    HIPCHK(c, hipSetDevice(c->device));
    const int nh = c->nharms;
    std::memset(out, 0, sizeof(double) * 2 * (size_t)(nh + 1) * (size_t)n);
    if (n == 0) return CS_OK;
    TabDev tb;
    if ((s = set_tables(c, r_lo, weights_f64 != 0, tb, t0 + n))) return s;
    if ((s = ensure(c, c->ftp, (size_t)nh * n * sizeof(double2)))) return s;
    GridDev g{t0, n, t0 + n};
    HIPCHK(c, launch_interp(c->st, spec_of(c), tb, g, t0, n, true, c->ftp.p, n));
    HIPCHK(c, hipMemcpyAsync(out + 2 * n, c->ftp.p, (size_t)nh * n * sizeof(double2),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

// Shared FP64 bridge: stack (already in c->ftp as double2, ld = n) -> fold k
// profiles -> sigma over nv columns (sample buffer `cap`) -> exact metrics.
static int bridge_metrics(cs_ctx* c, int k, int64_t n, int64_t nv, int64_t cap,
                          const std::vector<int>& ws, double scale, double* out) {
    const int nbins = 2 * (c->nharms / k);
    if (ws.empty() || ws.size() > 64) return c->fail(CS_EINVAL, "boxcar bank must hold 1..64 widths");
    for (int w : ws)
        if (w < 1 || w > nbins) return c->fail(CS_EINVAL, "boxcar widths must lie in 1..nbins");
    const double2* tw;
    int s;
    if ((s = get_twid(c, nbins, &tw))) return s;
    if ((s = ensure(c, c->prof, (size_t)nbins * n * sizeof(double)))) return s;
    if ((s = ensure(c, c->sigma, sizeof(double)))) return s;
    if ((s = ensure(c, c->mout, (size_t)std::max<int64_t>(nv, 1) * sizeof(double)))) return s;
    if ((s = upload(c, c->widths, std::vector<int32_t>(ws.begin(), ws.end())))) return s;
    double* prof = (double*)c->prof.p;
    HIPCHK(c, launch_c2r_f64(c->st, (const double2*)c->ftp.p, n, k, nbins, n, prof, n, tw, scale));
    HIPCHK(c, launch_sigma_f64(c->st, prof, n, nbins, nv, cap, (double*)c->sigma.p));
    HIPCHK(c, launch_exact_f64(c->st, prof, n, nbins, nv, (const int32_t*)c->widths.p, (int)ws.size(),
                               (const double*)c->sigma.p, (double*)c->mout.p));
    if (nv > 0)
        HIPCHK(c, hipMemcpyAsync(out, c->mout.p, (size_t)nv * sizeof(double), hipMemcpyDeviceToHost,
                                 c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

int cs_chunk_metrics(cs_ctx* c, double rstart, int64_t n, double lodr, int32_t weights_f64,
                     double* out) {
    (void)lodr;
    if (!c) return CS_EINVAL;
    if (n < 1) return CS_OK;
    int s = check_spectrum(c);
    if (s) return s;
    HIPCHK(c, hipSetDevice(c->device));
    const int nh = c->nharms, nbins = 2 * nh;
    TabDev tb;
    if ((s = set_tables(c, rstart, weights_f64 != 0, tb, n))) return s;
    if ((s = ensure(c, c->ftp, (size_t)nh * n * sizeof(double2)))) return s;
    GridDev g{0, n, n};
    HIPCHK(c, launch_interp(c->st, spec_of(c), tb, g, 0, n, true, c->ftp.p, n));
    int64_t cap = std::min<int64_t>((int64_t)nbins * n, kSigmaSamples);
    return bridge_metrics(c, 1, n, n, cap, ladder(c, nbins, 1), 1.0, out);
}

int cs_decim_metrics(cs_ctx* c, double rstart, int64_t n, int32_t k, int32_t weights_f64,
                     double* out, int64_t* nvalid) {
    if (!c || !nvalid) return CS_EINVAL;
    if (k < 1 || c->nharms / k < 1) return c->fail(CS_EINVAL, "invalid decimation factor");
    *nvalid = 0;
    if (n < 1) return CS_OK;
    int s = check_spectrum(c);
    if (s) return s;
    HIPCHK(c, hipSetDevice(c->device));
    const int nh = c->nharms, nbins = 2 * (nh / k);
    TabDev tb;
    if ((s = set_tables(c, rstart, weights_f64 != 0, tb, n))) return s;
    if ((s = ensure(c, c->ftp, (size_t)nh * n * sizeof(double2)))) return s;
    GridDev g{0, n, n};
    HIPCHK(c, launch_interp(c->st, spec_of(c), tb, g, 0, n, true, c->ftp.p, n));
    const double lodr = c->hidr / nh;
    int64_t nv = nvalid_of(k, rstart, lodr, n, (double)c->N / 2.0);
    *nvalid = nv;
    int64_t cap = std::min<int64_t>((int64_t)nbins * n, kSigmaSamples);
    return bridge_metrics(c, k, n, nv, cap, ladder(c, nbins, k), 1.0, out);
}

// The stacked harmonic amplitudes of reference_profiles (search.jl:1011-1051)
// for every fundamental of rfund into ftp[h-1][L]: per harmonic the row guard
// on the band, then FP64 Eqn. 30 point by point (kernel 1, :direct) or at the
// numbetween grid with the clamped linear interpolation (kernel 0, :fft).
static int block_interp(cs_ctx* c, const double* rfund, int64_t L, int32_t kernel) {
    if (kernel != 0 && kernel != 1) return c->fail(CS_EINVAL, "kernel must be :fft or :direct");
    int s = check_spectrum(c);
    if (s) return s;
    HIPCHK(c, hipSetDevice(c->device));
    const int nh = c->nharms, m2 = c->m / 2;
    std::vector<uint8_t> ok(nh, 0);
    std::vector<int64_t> lob(nh, 0), kgrid(nh, 1);
    const int64_t Nhalf = c->N / 2;
    for (int h = 1; h <= nh; ++h) {
        double rmin = rfund[0] * h, rmax = rfund[L - 1] * h;
        int64_t lobin = (int64_t)std::floor(rmin);
        int64_t hibin = (int64_t)std::ceil(rmax) + 1;
        int64_t numbins = hibin - lobin;
        ok[h - 1] = (lobin >= m2 && (lobin + numbins + m2) <= c->namps && hibin < Nhalf) ? 1 : 0;
        lob[h - 1] = lobin;
        kgrid[h - 1] = numbins * c->numbetween;
    }
    if (kernel == 0 && c->numbetween < 1) return c->fail(CS_EINVAL, "numbetween must be >= 1");
    if ((s = upload(c, c->rowok, ok))) return s;
    std::vector<double> rv(rfund, rfund + L);
    if ((s = upload(c, c->rfund, rv))) return s;
    if ((s = ensure(c, c->ftp, (size_t)nh * L * sizeof(double2)))) return s;
    if (kernel == 1) {
        HIPCHK(c, launch_point_interp_f64(c->st, spec_of(c), (const double*)c->rfund.p, L, nh, c->m,
                                          (const uint8_t*)c->rowok.p, (double2*)c->ftp.p, L));
    } else {
        std::vector<int64_t> tab(lob);
        tab.insert(tab.end(), kgrid.begin(), kgrid.end());
        if ((s = upload(c, c->bandtab, tab))) return s;
        const int64_t* dl = (const int64_t*)c->bandtab.p;
        HIPCHK(c, launch_point_interp_fft(c->st, spec_of(c), (const double*)c->rfund.p, L, nh, c->m,
                                          c->numbetween, dl, dl + nh, (const uint8_t*)c->rowok.p,
                                          (double2*)c->ftp.p, L));
    }
    return CS_OK;
}

int cs_block_metrics(cs_ctx* c, const double* rfund, int64_t L, int32_t kernel,
                     const int32_t* widths, int32_t nwidths, double* out) {
    if (!c) return CS_EINVAL;
    if (kernel != 0 && kernel != 1) return c->fail(CS_EINVAL, "kernel must be :fft or :direct");
    if (L < 1) return CS_OK;
    if (widths && nwidths < 1) return c->fail(CS_EINVAL, "empty width bank");
    int s = block_interp(c, rfund, L, kernel);
    if (s) return s;
    const int nbins = 2 * c->nharms;
    std::vector<int> ws;
    if (widths) ws.assign(widths, widths + nwidths);
    else ws = ladder(c, nbins, 1);
    int64_t cap = std::min<int64_t>((int64_t)nbins * L, kSigmaSamples);
    s = bridge_metrics(c, 1, L, L, cap, ws, 1.0 / (double)nbins, out);
    if (s) return s;
    for (int64_t j = 0; j < L; ++j)
        if (std::isnan(out[j])) return c->fail(CS_EDIM, "an interpolation window falls outside the amplitudes");
    return CS_OK;
}

int cs_reference_profiles(cs_ctx* c, const double* rfund, int64_t L, int32_t kernel, double* out) {
    if (!c || (L > 0 && (!rfund || !out))) return CS_EINVAL;
    if (kernel != 0 && kernel != 1) return c->fail(CS_EINVAL, "kernel must be :fft or :direct");
    if (L < 1) return CS_OK;
    int s = block_interp(c, rfund, L, kernel);
    if (s) return s;
    const int nbins = 2 * c->nharms;
    const double2* tw;
    if ((s = get_twid(c, nbins, &tw))) return s;
    if ((s = ensure(c, c->prof, (size_t)nbins * L * sizeof(double)))) return s;
    HIPCHK(c, launch_c2r_f64(c->st, (const double2*)c->ftp.p, L, 1, nbins, L, (double*)c->prof.p, L,
                             tw, 1.0 / (double)nbins));
    std::vector<double> pr((size_t)nbins * L);
    HIPCHK(c, hipMemcpyAsync(pr.data(), c->prof.p, pr.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int64_t j = 0; j < L; ++j)   // [bin][trial] -> column j of an nbins x L matrix
        for (int i = 0; i < nbins; ++i) out[j * nbins + i] = pr[(size_t)i * L + j];
    for (int64_t q = 0; q < (int64_t)nbins * L; ++q)
        if (std::isnan(out[q])) return c->fail(CS_EDIM, "an interpolation window falls outside the amplitudes");
    return CS_OK;
}

int cs_snr_metrics(cs_ctx* c, const double* profs, int32_t nbins, int64_t L, double fsp,
                   double maxfrac, int64_t sigma_samples, int32_t sigma_center,
                   const int32_t* widths, int32_t nwidths, double* out) {
    if (!c || (L > 0 && (!profs || !out))) return CS_EINVAL;
    if (nbins < 1 || nbins > 128) return c->fail(CS_EINVAL, "profiles of 1..128 bins");
    if (sigma_center != 0 && sigma_center != 1)
        return c->fail(CS_EINVAL, "sigma_center must be :zero or :median");
    if (L < 1) return CS_OK;
    std::vector<int> ws = widths ? std::vector<int>(widths, widths + nwidths)
                                 : boxcar_widths(nbins, fsp, maxfrac);
    if (ws.empty() || ws.size() > 64) return c->fail(CS_EINVAL, "boxcar bank must hold 1..64 widths");
    for (int w : ws)
        if (w < 1 || w > nbins) return c->fail(CS_EINVAL, "boxcar widths must lie in 1..nbins");
    HIPCHK(c, hipSetDevice(c->device));
    int s;
    std::vector<double> pr((size_t)nbins * L);   // column j of the matrix -> [bin][trial]
    for (int64_t j = 0; j < L; ++j)
        for (int i = 0; i < nbins; ++i) pr[(size_t)i * L + j] = profs[j * nbins + i];
    if ((s = upload(c, c->prof, pr))) return s;
    const int64_t cap = sigma_samples > 0 ? sigma_samples : (int64_t)nbins * L;
    const int64_t nsb = std::min<int64_t>((int64_t)nbins * L, cap);
    if ((s = ensure(c, c->mraw, (size_t)std::max<int64_t>(nsb, 1) * sizeof(double)))) return s;
    if ((s = ensure(c, c->sigma, sizeof(double)))) return s;
    if ((s = ensure(c, c->mout, (size_t)L * sizeof(double)))) return s;
    if ((s = upload(c, c->widths, std::vector<int32_t>(ws.begin(), ws.end())))) return s;
    HIPCHK(c, launch_sigma_any(c->st, (const double*)c->prof.p, L, nbins, L, cap, sigma_center,
                               (double*)c->mraw.p, (double*)c->sigma.p));
    HIPCHK(c, launch_exact_f64(c->st, (const double*)c->prof.p, L, nbins, L,
                               (const int32_t*)c->widths.p, (int)ws.size(),
                               (const double*)c->sigma.p, (double*)c->mout.p));
    HIPCHK(c, hipMemcpyAsync(out, c->mout.p, (size_t)L * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

int cs_snr_metrics_noalloc(cs_ctx* c, const double* d_profs, int32_t nbins, int64_t L,
                           double fsp, double maxfrac, int64_t sigma_samples,
                           int32_t sigma_center, const int32_t* widths, int32_t nwidths,
                           double* d_out) {
    if (!c || (L > 0 && (!d_profs || !d_out))) return CS_EINVAL;
    if (nbins < 1 || nbins > 128) return c->fail(CS_EINVAL, "profiles of 1..128 bins");
    if (sigma_center != 0 && sigma_center != 1)
        return c->fail(CS_EINVAL, "sigma_center must be :zero or :median");
    if (L < 1) return CS_OK;
    // the width bank, on the stack (boxcar_widths, search.jl:387-397, when none is given)
    WidthBank wb{};
    if (widths) {
        if (nwidths < 1 || nwidths > 64) return c->fail(CS_EINVAL, "boxcar bank must hold 1..64 widths");
        for (int a = 0; a < nwidths; ++a) wb.w[a] = widths[a];
        wb.n = nwidths;
    } else {
        const int wmax = std::max(1, (int)std::floor(maxfrac * nbins));
        for (int w = 1; w <= wmax; w = std::max((int)std::floor(fsp * w), w + 1)) {
            if (wb.n == 64) return c->fail(CS_EINVAL, "boxcar bank must hold 1..64 widths");
            wb.w[wb.n++] = w;
        }
    }
    for (int a = 0; a < wb.n; ++a)
        if (wb.w[a] < 1 || wb.w[a] > nbins)
            return c->fail(CS_EINVAL, "boxcar widths must lie in 1..nbins");
    const int64_t NN = (int64_t)nbins * L;
    const int64_t cap = sigma_samples > 0 ? sigma_samples : NN;
    if (std::min(NN, cap) > kScratchSamples)
        return c->fail(CS_EINVAL, "cs_snr_metrics_noalloc takes at most 2^20 sigma samples "
                                  "(min(nbins * L, sigma_samples)); cs_snr_metrics has no limit");
    if (!c->scratch_dev)
        return c->fail(CS_EDEVICE, std::string("no scratch address on this device: ") +
                                       hipGetErrorString(c->scratch_err));
    HIPCHK(c, hipSetDevice(c->device));
    unsigned char* base = (unsigned char*)c->scratch_dev;
    double* d_sigma = (double*)base;
    int32_t* d_w = (int32_t*)(base + 256);
    double* d_buf = (double*)(base + kScratchHead);
    // nothing in the scratch needs clearing: each kernel writes what the next reads
    hipLaunchKernelGGL(k_put_widths, dim3(1), dim3(64), 0, c->st, wb, d_w);
    HIPCHK(c, hipGetLastError());
    // profile j's bin i at d_profs[j * nbins + i]
    HIPCHK(c, launch_sigma_any(c->st, d_profs, 1, nbins, L, cap, sigma_center, d_buf, d_sigma,
                               nbins));
    HIPCHK(c, launch_exact_f64(c->st, d_profs, 1, nbins, L, d_w, wb.n, d_sigma, d_out, nbins));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

int cs_fourier_interp(cs_ctx* c, const double* r, int64_t n, int32_t m, double* out) {
    if (!c || (n > 0 && (!r || !out))) return CS_EINVAL;
    if (m < 2 || m % 2 != 0) return c->fail(CS_EINVAL, "m must be even");
    if (n < 1) return CS_OK;
    for (int64_t i = 0; i < n; ++i)
        if (!(r[i] >= 0.0)) return c->fail(CS_EINVAL, "r must be non-negative");
    int s = check_spectrum(c);
    if (s) return s;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<double> rv(r, r + n);
    if ((s = upload(c, c->rfund, rv))) return s;
    if ((s = upload(c, c->rowok, std::vector<uint8_t>(1, 1)))) return s;
    if ((s = ensure(c, c->ftp, (size_t)n * sizeof(double2)))) return s;
    HIPCHK(c, launch_point_interp_f64(c->st, spec_of(c), (const double*)c->rfund.p, n, 1, m,
                                      (const uint8_t*)c->rowok.p, (double2*)c->ftp.p, n));
    HIPCHK(c, hipMemcpyAsync(out, c->ftp.p, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int64_t i = 0; i < 2 * n; ++i)
        if (std::isnan(out[i])) return c->fail(CS_EDIM, "the interpolation window falls outside the amplitudes");
    return CS_OK;
}

// Candidate profiles and best widths (candidate.jl:37-61, 82-100).  H per
// candidate follows candidate_profile's stopping rule in FP64 on the host;
// skip[i] marks candidates measure_ducy leaves at NaN (H = -1 on the device).
static int cand_profiles(cs_ctx* c, const double* r, const int32_t* nharm, int64_t n, int32_t m,
                         bool ducy_rules, double* prof, int64_t ld, int32_t* len,
                         std::vector<int32_t>& bw) {
    if (m % 2 != 0 || m < 2) return c->fail(CS_EINVAL, "m must be even");
    int s = check_spectrum(c);
    if (s) return s;
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t m2 = m / 2, Nhalf = c->N / 2;
    std::vector<int32_t> H((size_t)n), woff((size_t)n + 1, 0), wtab;
    std::map<int, std::pair<int32_t, int32_t>> banks;   // profile bins -> (offset, count)
    for (int64_t i = 0; i < n; ++i) {
        if (ducy_rules ? !(nharm[i] >= 1 && r[i] > 0.0) : false) {
            H[i] = -1;
            woff[i + 1] = woff[i];
            continue;
        }
        if (nharm[i] < 1) return c->fail(CS_EINVAL, "nharm must be >= 1");
        int32_t h = 0;
        for (int32_t g = 1; g <= nharm[i]; ++g) {
            const double rh = r[i] * (double)g;
            const int64_t rint = (int64_t)std::floor(rh + 1e-15) + 1;
            if (!(rint - m2 >= 0 && rint + m2 <= c->namps && rh < (double)Nhalf)) break;
            h = g;
        }
        if (h > kMaxCandH)
            return c->fail(CS_EINVAL, "candidate profiles are limited to " +
                                          std::to_string(kMaxCandH) + " harmonics");
        H[i] = h;
        const int nb = h >= 1 ? 2 * h : 2;
        if (prof && ld < nb) return c->fail(CS_EDIM, "profile row stride too small");
        auto it = banks.find(nb);
        if (it == banks.end()) {
            std::vector<int> ws = boxcar_widths(nb, c->fsp, c->maxfrac);
            if (ws.empty() || ws.back() > nb)
                return c->fail(CS_EINVAL, "boxcar width exceeds the profile length");
            it = banks.emplace(nb, std::make_pair((int32_t)wtab.size(), (int32_t)ws.size())).first;
            wtab.insert(wtab.end(), ws.begin(), ws.end());
        }
        woff[i + 1] = woff[i] + it->second.second;
        if (len) len[i] = nb;
    }
    // per-candidate width lists (each candidate's bank copied in order)
    std::vector<int32_t> wcat;
    wcat.reserve((size_t)woff[n]);
    for (int64_t i = 0; i < n; ++i) {
        if (H[i] < 0) continue;
        const auto& b = banks[H[i] >= 1 ? 2 * H[i] : 2];
        wcat.insert(wcat.end(), wtab.begin() + b.first, wtab.begin() + b.first + b.second);
    }
    if (wcat.empty()) wcat.push_back(1);
    std::vector<double> rv(r, r + n);
    if ((s = upload(c, c->cand_r, rv))) return s;
    std::vector<int32_t> hw(H);
    hw.insert(hw.end(), woff.begin(), woff.end());
    hw.insert(hw.end(), wcat.begin(), wcat.end());
    if ((s = upload(c, c->cand_h, hw))) return s;
    if ((s = ensure(c, c->cand_w, (size_t)n * sizeof(int32_t)))) return s;
    if (prof && (s = ensure(c, c->cand_prof, (size_t)n * ld * sizeof(double)))) return s;
    const int32_t* hd = (const int32_t*)c->cand_h.p;
    HIPCHK(c, launch_cand_profiles(c->st, spec_of(c), (const double*)c->cand_r.p, hd, n, m,
                                   hd + 2 * n + 1, hd + n, prof ? (double*)c->cand_prof.p : nullptr,
                                   ld, (int32_t*)c->cand_w.p));
    bw.resize((size_t)n);
    HIPCHK(c, hipMemcpyAsync(bw.data(), c->cand_w.p, (size_t)n * sizeof(int32_t),
                             hipMemcpyDeviceToHost, c->st));
    if (prof)
        HIPCHK(c, hipMemcpyAsync(prof, c->cand_prof.p, (size_t)n * ld * sizeof(double),
                                 hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

int cs_candidate_profiles(cs_ctx* c, const double* r, const int32_t* nharm, int64_t n, int32_t m,
                          double* prof, int64_t ld, int32_t* len) {
    if (!c || (n > 0 && (!r || !nharm))) return CS_EINVAL;
    if (n <= 0) return CS_OK;
    std::vector<int32_t> bw;
    return cand_profiles(c, r, nharm, n, m, false, prof, ld, len, bw);
}

int cs_measure_ducy(cs_ctx* c, const double* r, const int32_t* nharm, int64_t n, int32_t m,
                    double* ducy) {
    if (!c || (n > 0 && (!r || !nharm || !ducy))) return CS_EINVAL;
    if (n <= 0) return CS_OK;
    std::vector<int32_t> len((size_t)n, 0), bw;
    int s = cand_profiles(c, r, nharm, n, m, true, nullptr, 0, len.data(), bw);
    if (s) return s;
    for (int64_t i = 0; i < n; ++i)
        ducy[i] = len[i] > 0 ? (double)bw[i] / (double)len[i] : NAN;
    return CS_OK;
}

int cs_plan_tables(const cs_params* p, double r_lo, int32_t h, int64_t* info, int32_t* row,
                   uint8_t* carry, float* W, float* A) {
    if (!p || p->m % 2 != 0 || p->nharms < 1 || h < 1 || h > p->nharms) return CS_EINVAL;
    Plans pl = build_plans(p->nharms, p->m, p->hidr, r_lo);
    if (!pl.rational) return CS_EINVAL;
    const HarmDev& d = pl.harm[h - 1];
    if (info) {
        info[0] = d.q; info[1] = d.s; info[2] = d.base_adv; info[3] = pl.pnum;
        info[4] = d.rfloor0; info[5] = d.P;
    }
    if (row) for (int i = 0; i < d.q; ++i) row[i] = pl.row[d.row_off + i];
    if (carry) for (int i = 0; i < d.P; ++i) carry[i] = pl.carry[d.tab_off + i];
    if (W) for (int i = 0; i < d.P * p->m; ++i) W[i] = pl.W32[(size_t)d.tab_off * p->m + i];
    if (A) for (int i = 0; i < 2 * d.P; ++i) A[i] = pl.A32[2 * (size_t)d.tab_off + i];
    return CS_OK;
}

// cs_search_files: every file's candidates go to one device buffer with a
// running count, the count after file `file` to ends[file]; the search
// neither resets that count nor waits for its stream
struct SweepSink {
    cs_cand_dev* cands;
    uint32_t* count;
    uint32_t cap;
    uint32_t* ends;
    int32_t file;
};

// cs_probe_split_bf16: the trial range whose stacked amplitudes the split-BF16
// kernel copies out, and where the amplitudes and bounds go on the host
struct ProbeReq {
    int64_t t0, n;
    float* amps;
    float* bounds;
};

// cs_search, or with a block mask (one byte per block, nonzero = search it)
// only the selected blocks: the batches are the runs of selected blocks
// Statistics request of cs_search_stats: histogram range and bins, windows.
struct StatsReq {
    double lo, hi;
    int nb, nwin;
};

// _logedges, search.jl:1406-1411: nwin log-spaced window edges over [a, b]
static std::vector<double> log_edges(double a, double b, int nwin) {
    a = std::max(a, std::numeric_limits<double>::min());
    b = std::max(b, a * (1 + 1e-9));
    const double la = std::log10(a), lb = std::log10(b);
    std::vector<double> e((size_t)nwin + 1);
    for (int i = 0; i <= nwin; ++i) e[i] = std::pow(10.0, la + (lb - la) * i / nwin);
    return e;
}

// _window_index, search.jl:1415-1416 (0-based): last edge <= f, clamped
static int window_index(const double* e, int nwin, double f) {
    int le = 0;
    while (le <= nwin && e[le] <= f) ++le;
    return std::min(std::max(le - 1, 0), nwin - 1);
}

// After a statistics search: windows of every (fold, block), the block
// records and the (fold, window) histograms, copied to the host.
// Every valid trial's exact metric into st_metric through the plain pipeline
// (FP32 interpolation, c2r, per-block sigma) and the per-trial exact scorer,
// in batches of the plain path's size.
static int stats_metrics_plain(cs_ctx* c, const FoldSet& fs0, const TabDev& tb, const SpecDev& sp,
                               const GridDev& g, int64_t total, int64_t Nprof, int64_t nchunks) {
    const int nf = fs0.nf, nh = c->nharms;
    const int64_t Bp = std::max<int64_t>(1, std::min<int64_t>(nchunks, kBatchTrials / Nprof));
    const int64_t BTp = Bp * Nprof;
    int64_t rows = 0;
    for (int f = 0; f < nf; ++f) rows += fs0.f[f].nbins;
    int s;
    if ((s = ensure(c, c->ftp, (size_t)nh * BTp * sizeof(float2)))) return s;
    if ((s = ensure(c, c->prof, (size_t)rows * BTp * sizeof(float)))) return s;
    FoldSet fs = fs0;
    fs.ldp = BTp;
    int64_t ro = 0;
    for (int f = 0; f < nf; ++f) {
        fs.prof[f] = (float*)c->prof.p + ro * BTp;
        ro += fs.f[f].nbins;
    }
    for (int64_t c0 = 0; c0 < nchunks; c0 += Bp) {
        const int64_t nb = std::min(Bp, nchunks - c0);
        const int64_t t_base = c0 * Nprof, nt = std::min(nb * Nprof, total - t_base);
        BatchDev b{};
        b.t_base = t_base;
        b.nt = nt;
        b.c_first = c0;
        b.nblocks = (int32_t)nb;
        b.nvalid = (const int32_t*)c->nvalid.p;
        b.sigma = (double*)c->sigma.p + c0 * nf;
        HIPCHK(c, launch_interp(c->st, sp, tb, g, t_base, nt, false, c->ftp.p, BTp));
        for (int f = 0; f < nf; ++f)
            HIPCHK(c, launch_c2r_f32(c->st, (const float2*)c->ftp.p, BTp, fs.f[f].k, fs.f[f].nbins,
                                     nt, fs.prof[f], BTp, fs.twid[f]));
        HIPCHK(c, launch_sigma_f32(c->st, fs, g, b));
        for (int f = 0; f < nf; ++f)
            HIPCHK(c, launch_exact_all(c->st, fs.prof[f], BTp, fs.f[f].nbins, fs.bankmask[f], f, nf,
                                       g, b, (double*)c->st_metric.p, total));
    }
    return CS_OK;
}

static int finish_stats(cs_ctx* c, const StatsReq& rq, const FoldSet& fs, const std::vector<int>& ks,
                        double r_lo, double r_hi, double lodr, int64_t total, int64_t Nprof,
                        int64_t nchunks) {
    auto& sh = c->sh;
    const int nf = fs.nf, nwin = rq.nwin, nb = rq.nb;
    const double T = c->T;
    sh.nf = nf;
    sh.nwin = nwin;
    sh.nb = nb;
    sh.lo = rq.lo;
    sh.hi = rq.hi;
    sh.r_lo = r_lo;
    sh.lodr = lodr;
    sh.T = T;
    sh.total = total;
    sh.Nprof = Nprof;
    sh.nchunks = nchunks;
    sh.N = c->N;
    sh.ks = ks;
    sh.Hks.assign(nf, 0);
    sh.edges.assign((size_t)nf * (nwin + 1), 0.0);
    sh.win.assign((size_t)nf * nchunks, 0);
    for (int f = 0; f < nf; ++f) {
        const int k = ks[f];
        sh.Hks[f] = fs.f[f].Hk;
        const std::vector<double> e = log_edges(k * r_lo / T, k * r_hi / T, nwin);
        std::copy(e.begin(), e.end(), sh.edges.begin() + (size_t)f * (nwin + 1));
        for (int64_t cc = 0; cc < nchunks; ++cc) {
            const int64_t i0 = cc * Nprof, n = std::min(Nprof, total - i0);
            const double rstart = r_lo + (double)i0 * lodr;
            const double rmean = rstart + (double)(n - 1) * lodr / 2;
            const double fb = k == 1 ? rmean / T : k * rmean / T;
            sh.win[(size_t)f * nchunks + cc] = window_index(e.data(), nwin, fb);
        }
    }
    int s;
    if ((s = upload(c, c->st_win, sh.win))) return s;
    if ((s = ensure(c, c->st_cnt, (size_t)nf * nwin * nb * sizeof(uint32_t)))) return s;
    if ((s = ensure(c, c->st_blk, (size_t)nf * nchunks * sizeof(BlockStatDev)))) return s;
    HIPCHK(c, hipMemsetAsync(c->st_cnt.p, 0, (size_t)nf * nwin * nb * sizeof(uint32_t), c->st));
    HIPCHK(c, launch_block_stats(c->st, (const double*)c->st_metric.p, total,
                                 (const int32_t*)c->nvalid.p, nf, Nprof, nchunks,
                                 (const int32_t*)c->st_win.p, nwin, rq.lo, rq.hi, nb,
                                 (uint32_t*)c->st_cnt.p, (BlockStatDev*)c->st_blk.p));
    sh.blk.resize((size_t)nf * nchunks);
    sh.cnt.resize((size_t)nf * nwin * nb);
    sh.nv.resize((size_t)nchunks * nf);
    HIPCHK(c, hipMemcpyAsync(sh.blk.data(), c->st_blk.p, sh.blk.size() * sizeof(BlockStatDev),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(sh.cnt.data(), c->st_cnt.p, sh.cnt.size() * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(sh.nv.data(), c->nvalid.p, sh.nv.size() * sizeof(int32_t),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    sh.valid = true;
    return CS_OK;
}

// Pinned result buffer of cs_search: the two counters (padded to kOutHead
// bytes), then room for the first kOutRecords candidates.
constexpr size_t kOutHead = 64;
constexpr int64_t kOutRecords = 1024;

static int search_impl(cs_ctx* c, double r_lo, double r_hi, double threshold, const uint8_t* mask,
                       int64_t nmask, cs_cand* out, int64_t cap, int64_t* n_out,
                       const StatsReq* sreq = nullptr, const ProbeReq* probe = nullptr,
                       const SweepSink* sweep = nullptr) {
    if (!c || !n_out) return CS_EINVAL;
    *n_out = 0;
    for (double& v : c->timings) v = 0.0;
    c->sig_blocks = 0;
    int s = check_spectrum(c);
    if (s) return s;
    HIPCHK(c, hipSetDevice(c->device));
    const int nh = c->nharms;
    const double lodr = c->hidr / nh;
    double span = std::floor((r_hi - r_lo) / lodr);
    int64_t total = span >= 0 ? (int64_t)span + 1 : 0;
    if (total <= 0) return CS_OK;
    const int64_t Nprof = c->Nprof;
    const int64_t nchunks = (total + Nprof - 1) / Nprof;
    TabDev tb;
    if ((s = set_tables(c, r_lo, false, tb, total))) return s;

    // folds: base pass, then every k > 1 in the order the parameters list them
    FoldSet fs;
    std::memset(&fs, 0, sizeof(fs));
    std::vector<int> ks{1};
    for (int k : c->decims) if (k > 1) ks.push_back(k);
    fs.nf = (int)ks.size();
    const int nf = fs.nf;
    int64_t rows_total = 0;
    for (int f = 0; f < nf; ++f) {
        FoldDev& fd = fs.f[f];
        fd.k = ks[f];
        fd.Hk = nh / ks[f];
        fd.nbins = 2 * fd.Hk;
        std::vector<int> ws = ladder(c, fd.nbins, fd.k);
        if ((int)ws.size() > kMaxWidths) return c->fail(CS_EINVAL, "boxcar bank too long");
        if (ws.back() > fd.nbins) return c->fail(CS_EINVAL, "boxcar width exceeds the profile length");
        fd.nw = (int)ws.size();
        for (int a = 0; a < fd.nw; ++a) fd.widths[a] = ws[a];
        fs.bankmask[f] = bank_mask(ws);
        fs.cap[f] = (int32_t)std::min<int64_t>((int64_t)fd.nbins * Nprof, kSigmaSamples);
        if ((s = get_twid(c, fd.nbins, &fs.twid[f]))) return s;
        rows_total += fd.nbins;
    }
    const bool mfma = c->mode == CS_MODE_FUSED;
    const bool fused = c->mode == CS_MODE_FUSED_VALU || mfma;
    if (fused && !(mfma ? mfma_supported(fs, nh) : fused_supported(fs, nh)))
        return c->fail(CS_EINVAL, "fused modes support profile lengths 16/20/24/30/32/40/60/64/120/128 "
                                  "with the default boxcar bank; use the plain mode");
    MfmaDev md{};
    int mfma_nwg = 0;
    if (mfma) {
        std::vector<int> nbs, wms, fks, key;
        for (int f = 0; f < nf; ++f) {
            nbs.push_back(fs.f[f].nbins);
            wms.push_back(fold_wm(fs.f[f].nbins, fs.bankmask[f]));
            fks.push_back(fs.f[f].k);
        }
        key = nbs;
        key.insert(key.end(), wms.begin(), wms.end());
        key.insert(key.end(), fks.begin(), fks.end());
        if (!(c->mfma_valid && c->mfma_rlo == r_lo && c->mfma_nbs == key)) {
            c->mfma_valid = false;
            MfmaTables mt = build_mfma_tables(build_plans(nh, c->m, c->hidr, r_lo), nbs, wms, fks);
            if (!mt.ok) return c->fail(CS_EINVAL, "fused MFMA mode: " + mt.why);
            if ((s = upload(c, c->m_whi, mt.wA_hi))) return s;
            if ((s = upload(c, c->m_wlo, mt.wA_lo))) return s;
            if ((s = upload(c, c->m_gA, mt.gA))) return s;
            if ((s = upload(c, c->m_hdesc, mt.hdesc))) return s;
            if ((s = upload(c, c->m_chi, mt.cA_hi))) return s;
            if ((s = upload(c, c->m_clo, mt.cA_lo))) return s;
            MfmaDev& mc = c->mfma;
            mc = MfmaDev{};
            mc.wA_hi = (const uint4*)c->m_whi.p;
            mc.wA_lo = (const uint4*)c->m_wlo.p;
            mc.gA = (const float2*)c->m_gA.p;
            mc.hdesc = (const int4*)c->m_hdesc.p;
            mc.cA_hi = (const uint4*)c->m_chi.p;
            mc.cA_lo = (const uint4*)c->m_clo.p;
            mc.S = mt.S;
            mc.U = mt.U;
            mc.q = mt.q;
            int ksmax = 1;
            for (int v : mt.ks) ksmax = std::max(ksmax, v);
            mc.ksmax = ksmax;
            for (int f = 0; f < kMaxFolds; ++f) {
                mc.c_base[f] = mt.c_base[f];
                mc.c_mt[f] = mt.c_mt[f];
                mc.c_ks[f] = mt.c_ks[f];
            }
            {   // sampled columns of a full block: eight int16 per block trial
                std::vector<int16_t> sc((size_t)Nprof * 8, (int16_t)-1);
                for (int f = 0; f < nf; ++f)
                    for (int64_t j = 0; j < Nprof; ++j)
                        sc[(size_t)j * 8 + f] =
                            (int16_t)mfma_sample_column(j, Nprof, fs.f[f].nbins, fs.cap[f]);
                if ((s = upload(c, c->m_scol, sc))) return s;
                mc.scol = (const uint4*)c->m_scol.p;
            }
            c->mfma_rlo = r_lo;
            c->mfma_nbs = key;
            c->mfma_valid = true;
        }
        md = c->mfma;
        if (total + 8 * md.S >= ((int64_t)1 << 32))
            return c->fail(CS_EINVAL, "more than 2^32 trials in one search");
        if (c->namps + 2 * kSpecPad >= ((int64_t)1 << 28))
            return c->fail(CS_EINVAL, "fused MFMA mode: spectra are limited to 2^28 amplitudes");
        if (Nprof > 65535 || Nprof < 16)
            return c->fail(CS_EINVAL, "fused MFMA mode: blocks of 16 to 65535 trials");
        if ((s = ensure(c, c->m_rowmask, (size_t)nchunks * sizeof(uint64_t)))) return s;
        const int64_t spl_ld = c->namps + 2 * kSpecPad;
        if (!c->spl_valid) {
            if ((s = ensure(c, c->spl, (size_t)2 * spl_ld * sizeof(uint32_t)))) return s;
            HIPCHK(c, launch_split_spectrum(c->st, spec_of(c), (uint32_t*)c->spl.p));
            c->spl_valid = true;
        }
        md.spl = (const uint32_t*)c->spl.p;
        md.rowmask = (const uint64_t*)c->m_rowmask.p;
        md.spl_ld = spl_ld;
        md.ntiles = (total + 8 * md.S - 1) / (8 * md.S) * md.U;
        md.tile_order = c->tune_order;
        md.generic = c->tune_generic;
        const MfmaLds lay = mfma_lds_plan(nh, fs, md);
        if (c->tune_noprefetch) md.wpf = 0;
        if (lay.total > 160 * 1024)
            return c->fail(CS_EINVAL, "fused MFMA mode: LDS plan exceeds 160 KB");
        // one persistent workgroup per compute unit (dealt over the 8 XCDs),
        // unless tuned
        mfma_nwg = c->tune_wg > 0 ? c->tune_wg : (c->ncu > 0 ? c->ncu : 256);
    }
    int64_t Bc =
        mfma ? std::min<int64_t>(nchunks, c->tune_batch > 0 ? c->tune_batch : kBatchBlocksMfma)
             : std::max<int64_t>(1, std::min<int64_t>(
                   nchunks, c->tune_batch > 0 ? c->tune_batch
                                              : (fused ? kBatchTrialsFused : kBatchTrials) / Nprof));
    if (mfma) {   // the split-BF16 kernel's 32-bit store offsets: samples and a fold's bounds
        const int64_t cap_s = (int64_t)kMfmaNoStore / ((int64_t)nf * kSigmaSamplesPadded * 4);
        const int64_t cap_m = (int64_t)kMfmaNoStore / ((int64_t)Nprof * 4);
        Bc = std::max<int64_t>(1, std::min(Bc, std::min(cap_s, cap_m)));
    }
    const int64_t BT = Bc * Nprof;
    fs.ldp = BT;

    // per-block valid counts for every fold (host, exact FP64 predicate)
    if (!(c->nv_valid && c->nv_rlo == r_lo && c->nv_total == total && c->nv_N == c->N)) {
        c->nv_valid = false;
        std::vector<int32_t> nv((size_t)nchunks * nf);
        const double nyq = (double)c->N / 2.0;
        for (int64_t cc = 0; cc < nchunks; ++cc) {
            int64_t i0 = cc * Nprof;
            int64_t n = std::min(Nprof, total - i0);
            double rstart = r_lo + (double)i0 * lodr;
            for (int f = 0; f < nf; ++f)
                nv[cc * nf + f] = (int32_t)nvalid_of(ks[f], rstart, lodr, n, nyq);
        }
        if ((s = upload(c, c->nvalid, nv))) return s;
        // the same counts as eight uint16 per block for the split-BF16 kernel
        // (one 16-byte load per trial; blocks of at most 65535 trials there)
        std::vector<uint16_t> nv16((size_t)nchunks * 8, 0);
        for (int64_t cc = 0; cc < nchunks; ++cc)
            for (int f = 0; f < nf; ++f) nv16[cc * 8 + f] = (uint16_t)std::min<int32_t>(nv[cc * nf + f], 65535);
        if ((s = upload(c, c->m_nv16, nv16))) return s;
        c->nv_rlo = r_lo;
        c->nv_total = total;
        c->nv_N = c->N;
        c->nv_valid = true;
    }
    if ((s = ensure(c, c->sigma, (size_t)nchunks * nf * sizeof(double)))) return s;
    FusedDev fu{};
    if (fused) {
        if ((s = ensure(c, c->mraw, (size_t)nf * BT * sizeof(float)))) return s;
        const int64_t sld = mfma ? kSigmaSamplesPadded : kSigmaSamples;
        if ((s = ensure(c, c->samples, (size_t)Bc * nf * sld * sizeof(float)))) return s;
        fu.sld = sld;
        fu.salign = mfma ? kSigmaSampleAlign : 1;
        if ((s = ensure(c, c->rowok, (size_t)Bc * nh))) return s;
        fu.mraw = (float*)c->mraw.p;
        fu.ldm = BT;
        if (probe) {
            if (!mfma) return c->fail(CS_EINVAL, "the split-BF16 probe needs CS_MODE_FUSED");
            if (Bc < nchunks)
                return c->fail(CS_EINVAL, "the split-BF16 probe covers one batch (at most " +
                                              std::to_string(Bc) + " blocks)");
            if (probe->t0 < 0 || probe->n < 1 || probe->t0 + probe->n > total)
                return c->fail(CS_EINVAL, "probe trials outside the band's " +
                                              std::to_string(total) + " trials");
            const size_t db = (size_t)nh * probe->n * sizeof(float2);
            if ((s = ensure(c, c->probe_buf, db))) return s;
            HIPCHK(c, hipMemsetAsync(c->probe_buf.p, 0, db, c->st));
            fu.dump = (float2*)c->probe_buf.p;
            fu.dump_t0 = probe->t0;
            fu.dump_n = probe->n;
        }
        fu.samples = (float*)c->samples.p;
        fu.rowok = (const uint8_t*)c->rowok.p;
    } else {
        if ((s = ensure(c, c->ftp, (size_t)nh * BT * sizeof(float2)))) return s;
        if ((s = ensure(c, c->prof, (size_t)rows_total * BT * sizeof(float)))) return s;
    }
    // plain pipeline: the c2r writes the sampled columns and the radix select
    // of the fused modes takes each block's median (generated c2r lengths);
    // otherwise the per-block select reads the profiles in place
    bool plain_samples = !fused;
    for (int f = 0; f < nf && plain_samples; ++f)
        plain_samples = c2r_samples_supported(fs.f[f].nbins);
    FusedDev fp{};
    if (plain_samples) {
        if ((s = ensure(c, c->samples, (size_t)Bc * nf * kSigmaSamples * sizeof(float)))) return s;
        fp.samples = (float*)c->samples.p;
        fp.sld = kSigmaSamples;
        fp.salign = 1;
    }
    if ((s = ensure(c, c->surv, (size_t)BT * nf * sizeof(uint64_t)))) return s;
    if ((s = ensure(c, c->counters, 4 * sizeof(uint32_t)))) return s;
    // precision :f64 (the reference's default profile stage): the sigma that
    // scales every reported metric comes from FP64 profiles
    const bool refine = c->precision == 0;
    RefineBufs rbufs{};
    if (refine) {
        const size_t npmax = (size_t)Bc * nf;
        if ((s = ensure(c, c->rflags, (npmax + 1) * sizeof(uint32_t)))) return s;
        if ((s = ensure(c, c->rlist, npmax * sizeof(uint32_t)))) return s;
        rbufs.cap = (uint32_t)std::min<size_t>(kRefineCap, npmax);
        if ((s = ensure(c, c->rsamp, (size_t)rbufs.cap * kSigmaSamples * sizeof(double)))) return s;
        rbufs.flags = (uint32_t*)c->rflags.p;
        rbufs.list = (uint32_t*)c->rlist.p;
        rbufs.samples = (double*)c->rsamp.p;
    }
    int64_t rowo = 0;
    for (int f = 0; f < nf && !fused; ++f) {
        fs.prof[f] = (float*)c->prof.p + rowo * BT;
        rowo += fs.f[f].nbins;
    }
    uint32_t cand_cap = (uint32_t)std::max<int64_t>(1 << 16, std::min<int64_t>(cap, (int64_t)1 << 26));
    if (mask && nmask != nchunks)
        return c->fail(CS_EDIM, "block mask has " + std::to_string(nmask) + " entries for " +
                                    std::to_string(nchunks) + " blocks");
    // batches: runs of consecutive selected blocks (first block, blocks);
    // on the split-BF16 path, groups of up to Bc selected blocks (first
    // selection position, blocks) with the tiles touching them
    std::vector<std::pair<int64_t, int64_t>> batches;
    std::vector<int32_t> selv;                 // selected blocks (mask given)
    std::vector<std::pair<int64_t, int64_t>> tile_rng;   // per batch: (list offset, tiles)
    std::vector<uint32_t> tiles;               // tile lists of a block subset
    if (mfma) {
        const int64_t sup = 8 * md.S;          // trials per super-tile (U tiles)
        if (mask) {
            for (int64_t cg = 0; cg < nchunks; ++cg)
                if (mask[cg]) selv.push_back((int32_t)cg);
        }
        const int64_t nsel = mask ? (int64_t)selv.size() : nchunks;
        for (int64_t p0 = 0; p0 < nsel; p0 += Bc) {
            const int64_t nb = std::min(Bc, nsel - p0);
            batches.emplace_back(p0, nb);
            if (!mask) {   // a run of blocks: one range of the identity list
                const int64_t lo = p0 * Nprof, hi = std::min((p0 + nb) * Nprof, total);
                const int64_t st_lo = lo / sup, st_hi = (hi + sup - 1) / sup;
                tile_rng.emplace_back(st_lo * md.U, (st_hi - st_lo) * md.U);
                continue;
            }
            const int64_t off = (int64_t)tiles.size();
            int64_t st_done = -1;              // super-tiles listed so far (ascending)
            for (int64_t p = p0; p < p0 + nb; ++p) {
                const int64_t lo = (int64_t)selv[p] * Nprof;
                const int64_t hi = std::min(lo + Nprof, total);
                for (int64_t st = std::max(lo / sup, st_done + 1); st <= (hi - 1) / sup; ++st) {
                    for (int u = 0; u < md.U; ++u) tiles.push_back((uint32_t)(st * md.U + u));
                    st_done = st;
                }
            }
            tile_rng.emplace_back(off, (int64_t)tiles.size() - off);
        }
    } else if (mask && !fused) {
        // plain pipeline over selected blocks: lists of up to Bc blocks
        for (int64_t cg = 0; cg < nchunks; ++cg)
            if (mask[cg]) selv.push_back((int32_t)cg);
        for (int64_t p0 = 0; p0 < (int64_t)selv.size(); p0 += Bc)
            batches.emplace_back(p0, std::min<int64_t>(Bc, (int64_t)selv.size() - p0));
    } else {
        for (int64_t c0 = 0; c0 < nchunks;) {
            if (mask && !mask[c0]) {
                ++c0;
                continue;
            }
            int64_t e = c0;
            while (e < nchunks && (!mask || mask[e]) && e - c0 < Bc) ++e;
            batches.emplace_back(c0, e - c0);
            c0 = e;
        }
    }
    const int64_t nbatch = (int64_t)batches.size();
    const int nev = 6;
    while ((int64_t)c->ev.size() < nbatch * nev) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->ev.push_back(e);
    }
    GridDev g{0, Nprof, total};
    SpecDev sp = spec_of(c);
    uint32_t* surv_count = (uint32_t*)c->counters.p;
    uint32_t* cand_count = surv_count + 1;
    double medcut = (c->medmargin == INFINITY) ? -INFINITY : threshold - c->medmargin;
    // statistics: every valid metric exact and kept, [fold][trial]
    struct StatsGuard {
        cs_ctx* c;
        ~StatsGuard() { c->stats_on = false; }
    } sguard{c};
    // The metrics come from FP32 profiles scored exactly per trial before the
    // search (every fold of a supported length and default-bank widths), which
    // then select the search's survivors; or else from the exact rescan of
    // every trial with the gate disabled.  Either way the candidates are those
    // of the exact scan everywhere, as when the reference collects statistics.
    bool stats_fast = sreq != nullptr;
    for (int f = 0; f < nf && stats_fast; ++f)
        stats_fast = exact_all_supported(fs.f[f].nbins) && fs.bankmask[f] != 0;
    if (sreq) {
        c->sh.valid = false;
        if ((s = ensure(c, c->st_metric, (size_t)nf * total * sizeof(double)))) return s;
        HIPCHK(c, hipMemsetAsync(c->st_metric.p, 0xFF, (size_t)nf * total * sizeof(double), c->st));
        if (!stats_fast) {
            medcut = -INFINITY;
            c->stats_on = true;
        }
    }

    const bool plain_list = mask && !fused;
    if (plain_list) {
        if ((s = upload(c, c->m_sel, selv))) return s;
        if ((s = ensure(c, c->m_slot, (size_t)nchunks * sizeof(int32_t)))) return s;
        HIPCHK(c, launch_slot_table(c->st, (const int32_t*)c->m_sel.p, (int64_t)selv.size(),
                                    nchunks, (int32_t*)c->m_slot.p));
    }
    if (mfma) {
        // the fold and table descriptors, once per search (the batches only
        // differ in their by-value BatchDev)
        // selection tables: the identity list serves every block selected
        const int64_t nid = std::max(nchunks, md.ntiles);
        if (c->m_ident_n < nid) {
            c->m_ident_n = 0;
            if ((s = ensure(c, c->m_ident, (size_t)nid * sizeof(uint32_t)))) return s;
            HIPCHK(c, launch_iota(c->st, (uint32_t*)c->m_ident.p, nid));
            c->m_ident_n = nid;
        }
        if (mask) {
            if ((s = upload(c, c->m_sel, selv))) return s;
            if ((s = upload(c, c->m_tiles, tiles))) return s;
            if ((s = ensure(c, c->m_slot, (size_t)nchunks * sizeof(int32_t)))) return s;
            HIPCHK(c, launch_slot_table(c->st, (const int32_t*)c->m_sel.p, (int64_t)selv.size(),
                                        nchunks, (int32_t*)c->m_slot.p));
        }
        md.nv16 = (const uint4*)c->m_nv16.p;
        std::vector<uint8_t> desc(sizeof(FoldSet) + sizeof(MfmaDev));
        std::memcpy(desc.data(), &fs, sizeof(FoldSet));
        std::memcpy(desc.data() + sizeof(FoldSet), &md, sizeof(MfmaDev));
        // the same descriptors as the last search need no upload (and no sync)
        if (!c->m_desc.p || desc != c->m_desc_last) {
            c->m_desc_last.clear();
            if ((s = upload(c, c->m_desc, desc))) return s;
            c->m_desc_last = desc;
        }
    }
    // The launch sequence of one attempt: the row masks, the counters, then
    // every batch's kernels; stage events only when not captured into a graph.
    auto emit = [&](bool ev) -> int {
        if (mfma) HIPCHK(c, launch_rowmask(c->st, sp, tb, g, nchunks, (uint64_t*)c->m_rowmask.p));
        // the counters, and the first batch's refinement flags and count,
        // in one launch
        const int rpair0 = (refine && nbatch > 0) ? (int)(batches[0].second * nf) + 1 : 0;
        HIPCHK(c, launch_reset2_u32(c->st, surv_count, sweep ? 1 : 2,
                                    rpair0 ? rbufs.flags : nullptr, rpair0));
        for (int64_t bi = 0; bi < nbatch; ++bi) {
            hipEvent_t* E = &c->ev[bi * nev];
            auto rec = [&](int q) { return ev ? hipEventRecord(E[q], c->st) : hipSuccess; };
            const int64_t c0 = batches[bi].first;
            const int64_t nb = batches[bi].second;
            const int64_t t_base = c0 * Nprof;
            const int64_t nt = std::min(nb * Nprof, total - t_base);
            BatchDev b{};
            b.t_base = t_base;
            b.nt = nt;
            b.c_first = c0;
            b.nblocks = (int32_t)nb;
            b.nvalid = (const int32_t*)c->nvalid.p;
            b.sigma = (double*)c->sigma.p + c0 * nf;
            if (plain_list) {   // c0 is a selection position; columns block by block
                b.t_base = b.c_first = 0;
                b.nt = nb * Nprof;
                b.slot0 = (int32_t)c0;
                b.sel = (const int32_t*)c->m_sel.p;
                b.slot = (const int32_t*)c->m_slot.p;
            }
            if (mfma) {   // c0 is a selection position
                b.t_base = b.nt = b.c_first = 0;
                b.slot0 = (int32_t)c0;
                const uint32_t* ident = (const uint32_t*)c->m_ident.p;
                b.sel = mask ? (const int32_t*)c->m_sel.p : (const int32_t*)ident;
                b.slot = mask ? (const int32_t*)c->m_slot.p : (const int32_t*)ident;
                b.tiles = (mask ? (const uint32_t*)c->m_tiles.p : ident) + tile_rng[bi].first;
                b.ntile = tile_rng[bi].second;
            }
            OutDev o;
            o.surv = (uint64_t*)c->surv.p;
            o.surv_count = surv_count;
            o.surv_cap = (uint32_t)(BT * nf);
            o.cand_cap = sweep ? sweep->cap : cand_cap;
            o.cands = sweep ? sweep->cands : (cs_cand_dev*)c->cands.p;
            o.cand_count = sweep ? sweep->count : cand_count;
            o.mall = c->stats_on ? (double*)c->st_metric.p : nullptr;
            o.mall_ld = total;
            // statistics fast path: the every-trial FP32 exact metrics select
            // the survivors (everything within 0.01 of the threshold, far
            // beyond their 1e-5 relative difference from this mode's FP64
            // metrics), so the candidates are the exact-everywhere ones
            o.gm = (sreq && stats_fast) ? (const double*)c->st_metric.p : nullptr;
            o.gm_ld = total;
            o.gmcut = threshold - 0.01;
            o.threshold = threshold;
            o.medcut = medcut;
            HIPCHK(c, rec(0));
            if (fused) {
                if (!mfma) HIPCHK(c, launch_guard(c->st, sp, tb, g, b, (uint8_t*)c->rowok.p));
                if (mfma) {
                    const int nwg = (int)std::min<int64_t>(mfma_nwg, b.ntile);
                    HIPCHK(c, launch_fused_mfma(c->st, sp, tb, g, fs, b, fu, md,
                                                (const FoldSet*)c->m_desc.p,
                                                (const MfmaDev*)((const uint8_t*)c->m_desc.p + sizeof(FoldSet)),
                                                nwg));
                }
                else HIPCHK(c, launch_fused(c->st, sp, tb, g, fs, b, fu));
                HIPCHK(c, rec(1));
                HIPCHK(c, rec(2));
                // (the first batch's count is zero since the reset above:
                // nothing between pushes survivors)
                if (bi > 0) HIPCHK(c, launch_reset_u32(c->st, surv_count, 1));
                HIPCHK(c, launch_select(c->st, fs, g, b, fu, o));
                HIPCHK(c, rec(3));
                HIPCHK(c, rec(4));
            } else {
                HIPCHK(c, launch_interp(c->st, sp, tb, g, b.t_base, b.nt, false, c->ftp.p, BT, b.sel,
                                        b.slot0));
                HIPCHK(c, rec(1));
                for (int f = 0; f < nf; ++f) {
                    if (plain_samples) {
                        SampleSink sk{fp.samples, fp.sld, b.nvalid, g, b.t_base, b.c_first, f, nf,
                                      fs.cap[f], b.slot0, b.sel};
                        HIPCHK(c, launch_c2r_samples(c->st, (const float2*)c->ftp.p, BT,
                                                     fs.f[f].k, fs.f[f].nbins, b.nt, fs.prof[f],
                                                     BT, sk));
                    } else {
                        HIPCHK(c, launch_c2r_f32(c->st, (const float2*)c->ftp.p, BT, fs.f[f].k,
                                                 fs.f[f].nbins, b.nt, fs.prof[f], BT, fs.twid[f]));
                    }
                }
                HIPCHK(c, rec(2));
                if (plain_samples) HIPCHK(c, launch_select(c->st, fs, g, b, fp, o));
                else HIPCHK(c, launch_sigma_f32(c->st, fs, g, b));
                HIPCHK(c, rec(3));
                HIPCHK(c, launch_reset_u32(c->st, surv_count, 1));
                HIPCHK(c, launch_gate_f32(c->st, fs, g, b, o));
                HIPCHK(c, rec(4));
            }
            if (refine) {
                RefineBufs rb = rbufs;
                rb.flags_clear = bi == 0;
                HIPCHK(c, launch_sigma_refine(c->st, sp, tb, g, fs, b, o, rb));
            }
            HIPCHK(c, launch_rescan(c->st, sp, tb, g, fs, b, o, 8192));
            HIPCHK(c, rec(5));
        }
        return CS_OK;
    };
    const bool graph = c->use_graph && !sreq && !mask && !probe && !sweep;
    std::vector<uint8_t> gkey;
    if (graph) {   // everything the captured launches bake in
        auto put = [&](const void* p, size_t n) {
            const uint8_t* b = (const uint8_t*)p;
            gkey.insert(gkey.end(), b, b + n);
        };
        const void* ptrs[] = {c->surv.p, c->counters.p, c->nvalid.p, c->sigma.p, c->m_rowmask.p,
                              c->m_desc.p, c->m_ident.p, c->ftp.p, c->prof.p, c->rowok.p,
                              c->mraw.p, c->samples.p, refine ? c->rflags.p : nullptr,
                              refine ? c->rlist.p : nullptr, refine ? c->rsamp.p : nullptr};
        put(ptrs, sizeof(ptrs));
        put(&sp, sizeof(sp));
        put(&tb, sizeof(tb));
        put(&g, sizeof(g));
        put(&fs, sizeof(fs));
        put(&md, sizeof(md));
        put(&fu, sizeof(fu));
        const double dv[] = {threshold, medcut};
        put(dv, sizeof(dv));
        const int64_t iv[] = {c->mode, mfma_nwg, nbatch, BT, nchunks, total, Nprof};
        put(iv, sizeof(iv));
        put(batches.data(), batches.size() * sizeof(batches[0]));
        put(tile_rng.data(), tile_rng.size() * sizeof(tile_rng[0]));
    }
    if (sreq && stats_fast && (s = stats_metrics_plain(c, fs, tb, sp, g, total, Nprof, nchunks)))
        return s;
    if (sweep) {   // the caller reads every file's results after its last search
        c->sig_blocks = 0;
        if ((s = emit(false))) return s;
        HIPCHK(c, launch_copy_u32(c->st, sweep->count, sweep->ends + sweep->file));
        return CS_OK;
    }
    for (int attempt = 0; attempt < 2; ++attempt) {
        if ((s = ensure(c, c->cands, (size_t)cand_cap * sizeof(cs_cand_dev)))) return s;
        if (graph) {
            std::vector<uint8_t> key = gkey;
            const void* ck[] = {c->cands.p};
            key.insert(key.end(), (const uint8_t*)ck, (const uint8_t*)ck + sizeof(ck));
            key.insert(key.end(), (const uint8_t*)&cand_cap, (const uint8_t*)&cand_cap + sizeof(cand_cap));
            if (!c->g_exec || key != c->g_key) {
                if (c->g_exec) (void)hipGraphExecDestroy(c->g_exec);
                if (c->g_graph) (void)hipGraphDestroy(c->g_graph);
                c->g_exec = nullptr;
                c->g_graph = nullptr;
                c->g_key.clear();
                HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed));
                const int se = emit(false);
                hipGraph_t gr = nullptr;
                const hipError_t ee = hipStreamEndCapture(c->st, &gr);
                if (se) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return se;
                }
                HIPCHK(c, ee);
                c->g_graph = gr;
                HIPCHK(c, hipGraphInstantiate(&c->g_exec, gr, nullptr, nullptr, 0));
                c->g_key = key;
                ++c->g_captures;
            }
            HIPCHK(c, hipEventRecord(c->ev[0], c->st));
            HIPCHK(c, hipGraphLaunch(c->g_exec, c->st));
            HIPCHK(c, hipEventRecord(c->ev[1], c->st));
        } else if ((s = emit(true))) {
            return s;
        }
        // the counters and the first candidates come back in one synchronisation
        // (queued behind the search on its stream, into pinned memory)
        if (!c->hout && hipHostMalloc(&c->hout, kOutHead + kOutRecords * sizeof(cs_cand_dev),
                                      hipHostMallocDefault) != hipSuccess) {
            c->hout = nullptr;
            return c->fail(CS_ENOMEM, "pinned result buffer allocation failed");
        }
        const int64_t first = std::min<int64_t>(cand_cap, kOutRecords);
        HIPCHK(c, hipMemcpyAsync(c->hout, c->counters.p, 2 * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync((char*)c->hout + kOutHead, c->cands.p,
                                 (size_t)first * sizeof(cs_cand_dev), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        uint32_t counts[2];
        std::memcpy(counts, c->hout, sizeof(counts));
        if (counts[1] > cand_cap && attempt == 0) {
            cand_cap = counts[1] + 1024;
            continue;
        }
        int64_t nc = std::min<int64_t>(counts[1], cand_cap);
        std::vector<cs_cand_dev> hc((size_t)nc);
        const int64_t n0 = std::min<int64_t>(nc, first);
        if (n0)
            std::memcpy(hc.data(), (const char*)c->hout + kOutHead, (size_t)n0 * sizeof(cs_cand_dev));
        if (nc > n0)
            HIPCHK(c, hipMemcpy(hc.data() + n0, (const cs_cand_dev*)c->cands.p + n0,
                                (size_t)(nc - n0) * sizeof(cs_cand_dev), hipMemcpyDeviceToHost));
        std::sort(hc.begin(), hc.end(), [](const cs_cand_dev& a, const cs_cand_dev& b) {
            return a.t != b.t ? a.t < b.t : a.k < b.k;
        });
        if (graph) {   // one interval: the whole replayed sequence
            float tot = 0.f;
            HIPCHK(c, hipEventElapsedTime(&tot, c->ev[0], c->ev[1]));
            c->timings[6] = tot;
        }
        for (int64_t bi = 0; bi < nbatch && !graph; ++bi) {
            hipEvent_t* E = &c->ev[bi * nev];
            for (int q = 0; q < 5; ++q) {
                float ms = 0.f;
                HIPCHK(c, hipEventElapsedTime(&ms, E[q], E[q + 1]));
                // fused mode: interval 0 is the fused interpolation/c2r/bound kernel
                c->timings[(fused && q == 0) ? 5 : q] += ms;
            }
            float tot = 0.f;
            HIPCHK(c, hipEventElapsedTime(&tot, E[0], E[5]));
            c->timings[6] += tot;
        }
        c->timings[7] = (double)nbatch;
        if (sreq && (s = finish_stats(c, *sreq, fs, ks, r_lo, r_hi, lodr, total, Nprof, nchunks)))
            return s;
        c->sig_blocks = nchunks;
        c->sig_nf = nf;
        c->sig_sparse = mask != nullptr;
        if (probe) {   // one batch: mraw column = global trial
            if (probe->amps)
                HIPCHK(c, hipMemcpy(probe->amps, c->probe_buf.p,
                                    (size_t)nh * probe->n * sizeof(float2), hipMemcpyDeviceToHost));
            for (int f = 0; f < nf && probe->bounds; ++f)
                HIPCHK(c, hipMemcpy(probe->bounds + (size_t)f * probe->n,
                                    (const float*)c->mraw.p + (size_t)f * BT + probe->t0,
                                    (size_t)probe->n * sizeof(float), hipMemcpyDeviceToHost));
        }
        *n_out = nc;
        if (out) {
            int64_t w = std::min<int64_t>(nc, cap);
            for (int64_t i = 0; i < w; ++i) {
                out[i].t = hc[i].t;
                out[i].k = hc[i].k;
                out[i].nharm = hc[i].nharm;
                out[i].block = hc[i].block;
                out[i].metric = hc[i].metric;
            }
        }
        if (nc > cap) return c->fail(CS_ENOSPC, "candidate buffer too small");
        return CS_OK;
    }
    return c->fail(CS_ENOSPC, "candidate buffer could not be sized");
}

int cs_probe_split_bf16(cs_ctx* c, double r_lo, double r_hi, int64_t t0, int64_t n, float* amps,
                        float* bounds) {
    if (!c) return CS_EINVAL;
    if (c->mode != CS_MODE_FUSED) return c->fail(CS_EINVAL, "the split-BF16 probe needs CS_MODE_FUSED");
    ProbeReq pr{t0, n, amps, bounds};
    int64_t nout = 0;
    return search_impl(c, r_lo, r_hi, INFINITY, nullptr, 0, nullptr, 0, &nout, nullptr, &pr);
}

int cs_search(cs_ctx* c, double r_lo, double r_hi, double threshold, cs_cand* out, int64_t cap,
              int64_t* n_out) {
    return search_impl(c, r_lo, r_hi, threshold, nullptr, 0, out, cap, n_out);
}

int cs_search_blocks(cs_ctx* c, double r_lo, double r_hi, double threshold, const uint8_t* mask,
                     int64_t nmask, cs_cand* out, int64_t cap, int64_t* n_out) {
    if (!c || !mask) return CS_EINVAL;
    return search_impl(c, r_lo, r_hi, threshold, mask, nmask, out, cap, n_out);
}

// ---- metric statistics and the normalised pass --------------------------------
int cs_search_stats(cs_ctx* c, double r_lo, double r_hi, double threshold, const cs_stats_cfg* cfg,
                    cs_cand* out, int64_t cap, int64_t* n_out) {
    if (!c || !cfg || !n_out) return CS_EINVAL;
    if (!(cfg->hist_hi > cfg->hist_lo))
        return c->fail(CS_EINVAL, "histogram hi must exceed lo");
    if (cfg->hist_nb < 1) return c->fail(CS_EINVAL, "histogram bin count must be >= 1");
    if (cfg->hist_nb > kStatsMaxBins)
        return c->fail(CS_EINVAL, "histograms are limited to " + std::to_string(kStatsMaxBins) + " bins");
    StatsReq rq{cfg->hist_lo, cfg->hist_hi, cfg->hist_nb, std::max(1, (int)cfg->nwin)};
    c->sh.valid = false;
    return search_impl(c, r_lo, r_hi, threshold, nullptr, 0, out, cap, n_out, &rq);
}

// folds of the last statistics search in ascending k (the reported order)
static std::vector<int> folds_by_k(const cs_ctx* c) {
    std::vector<int> o(c->sh.nf);
    for (int f = 0; f < c->sh.nf; ++f) o[f] = f;
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return c->sh.ks[a] < c->sh.ks[b]; });
    return o;
}

int cs_stats_blocks(cs_ctx* c, cs_block_stat* out, int64_t cap, int64_t* n) {
    if (!c || !n) return CS_EINVAL;
    const auto& sh = c->sh;
    if (!sh.valid) return c->fail(CS_EINVAL, "no statistics search has run");
    const std::vector<int> ord = folds_by_k(c);
    const double nyq2 = (double)sh.N / 2;
    int64_t w = 0;
    for (int64_t cc = 0; cc < sh.nchunks; ++cc) {
        const int64_t i0 = cc * sh.Nprof, nblk = std::min(sh.Nprof, sh.total - i0);
        const double rstart = sh.r_lo + (double)i0 * sh.lodr;
        const double rmean = rstart + (double)(nblk - 1) * sh.lodr / 2;
        for (int f : ord) {
            const BlockStatDev& b = sh.blk[(size_t)f * sh.nchunks + cc];
            const int k = sh.ks[f], Hk = sh.Hks[f];
            if (b.n == 0) continue;   // no valid trial (k > 1 past Nyquist): no record
            if (out && w < cap) {
                cs_block_stat& r = out[w];
                r.block = cc;
                r.k = k;
                r.Hk = Hk;
                r.nbins = 2 * Hk;
                r.pad = 0;
                r.n = b.n;
                // chunk_ngoodbins (search.jl:206-207) at k rmean
                r.ngoodbins = std::min(nyq2 / (k == 1 ? rmean : k * rmean), (double)Hk);
                if (k == 1) {
                    r.flo = rstart / sh.T;
                    r.fhi = (rstart + (double)(nblk - 1) * sh.lodr) / sh.T;
                } else {
                    r.flo = k * rstart / sh.T;
                    r.fhi = k * (rstart + (double)(b.n - 1) * sh.lodr) / sh.T;
                }
                r.min = b.vmin;
                r.median = b.median;
                r.mean = b.mean;
                r.std = b.stdv;
                r.max = b.vmax;
            }
            ++w;
        }
    }
    *n = w;
    if (out && w > cap) return c->fail(CS_ENOSPC, "block statistics buffer too small");
    return CS_OK;
}

int cs_stats_hists(cs_ctx* c, int64_t* counts, cs_hist_acc* acc, int32_t* nfolds) {
    if (!c) return CS_EINVAL;
    const auto& sh = c->sh;
    if (!sh.valid) return c->fail(CS_EINVAL, "no statistics search has run");
    if (nfolds) *nfolds = sh.nf;
    const std::vector<int> ord = folds_by_k(c);
    for (int o = 0; o < sh.nf; ++o) {
        const int f = ord[o];
        const double* e = sh.edges.data() + (size_t)f * (sh.nwin + 1);
        for (int w = 0; w < sh.nwin; ++w) {
            if (counts)
                for (int i = 0; i < sh.nb; ++i)
                    counts[((size_t)o * sh.nwin + w) * sh.nb + i] =
                        sh.cnt[((size_t)f * sh.nwin + w) * sh.nb + i];
            if (!acc) continue;
            cs_hist_acc& a = acc[(size_t)o * sh.nwin + w];
            a.k = sh.ks[f];
            a.win = w + 1;
            a.under = a.over = a.total = 0;
            a.flo = e[w];
            a.fhi = e[w + 1];
            a.sum = a.sumsq = 0.0;
            a.vmin = INFINITY;
            a.vmax = -INFINITY;
            for (int64_t cc = 0; cc < sh.nchunks; ++cc) {
                if (sh.win[(size_t)f * sh.nchunks + cc] != w) continue;
                const BlockStatDev& b = sh.blk[(size_t)f * sh.nchunks + cc];
                if (b.n == 0) continue;
                a.under += b.under;
                a.over += b.over;
                a.total += b.n;
                a.sum += b.sum;
                a.sumsq += b.sumsq;
                a.vmin = std::min(a.vmin, b.vmin);
                a.vmax = std::max(a.vmax, b.vmax);
            }
        }
    }
    return CS_OK;
}

int cs_stats_metrics(cs_ctx* c, int32_t k, double* out, int64_t n) {
    if (!c || !out) return CS_EINVAL;
    const auto& sh = c->sh;
    if (!sh.valid) return c->fail(CS_EINVAL, "no statistics search has run");
    if (n != sh.total) return c->fail(CS_EDIM, "metric buffer must hold every trial of the search");
    for (int f = 0; f < sh.nf; ++f)
        if (sh.ks[f] == k) {
            HIPCHK(c, hipSetDevice(c->device));
            HIPCHK(c, hipMemcpy(out, (const double*)c->st_metric.p + (size_t)f * sh.total,
                                (size_t)sh.total * sizeof(double), hipMemcpyDeviceToHost));
            return CS_OK;
        }
    return c->fail(CS_EINVAL, "k is not a searched decimation");
}

int cs_search_normalized(cs_ctx* c, double threshold, const double* loc, const double* scale,
                         cs_cand* out, int64_t cap, int64_t* n_out) {
    if (!c || !loc || !scale || !n_out) return CS_EINVAL;
    *n_out = 0;
    const auto& sh = c->sh;
    if (!sh.valid) return c->fail(CS_EINVAL, "no statistics search has run");
    HIPCHK(c, hipSetDevice(c->device));
    const std::vector<int> ord = folds_by_k(c);
    std::vector<double> tab((size_t)sh.nf * (sh.nwin + 1) + 2 * (size_t)sh.nf * sh.nwin);
    double* ed = tab.data();
    double* lc = ed + (size_t)sh.nf * (sh.nwin + 1);
    double* sc = lc + (size_t)sh.nf * sh.nwin;
    std::copy(sh.edges.begin(), sh.edges.end(), ed);
    for (int o = 0; o < sh.nf; ++o)
        for (int w = 0; w < sh.nwin; ++w) {
            lc[(size_t)ord[o] * sh.nwin + w] = loc[(size_t)o * sh.nwin + w];
            sc[(size_t)ord[o] * sh.nwin + w] = scale[(size_t)o * sh.nwin + w];
        }
    int s;
    if ((s = upload(c, c->st_norm, tab))) return s;
    NormDev nd{};
    nd.metrics = (const double*)c->st_metric.p;
    nd.ld = sh.total;
    nd.total = sh.total;
    nd.Nprof = sh.Nprof;
    nd.r_lo = sh.r_lo;
    nd.lodr = sh.lodr;
    nd.T = sh.T;
    nd.threshold = threshold;
    nd.edges = (const double*)c->st_norm.p;
    nd.loc = nd.edges + (size_t)sh.nf * (sh.nwin + 1);
    nd.scale = nd.loc + (size_t)sh.nf * sh.nwin;
    nd.nwin = sh.nwin;
    nd.nf = sh.nf;
    for (int f = 0; f < sh.nf; ++f) {
        nd.k[f] = sh.ks[f];
        nd.Hk[f] = sh.Hks[f];
    }
    uint32_t cand_cap = (uint32_t)std::max<int64_t>(1 << 16, std::min<int64_t>(cap, (int64_t)1 << 26));
    if ((s = ensure(c, c->counters, 4 * sizeof(uint32_t)))) return s;
    for (int attempt = 0; attempt < 2; ++attempt) {
        if ((s = ensure(c, c->cands, (size_t)cand_cap * sizeof(cs_cand_dev)))) return s;
        uint32_t* cnt = (uint32_t*)c->counters.p + 1;
        HIPCHK(c, launch_reset_u32(c->st, cnt, 1));
        HIPCHK(c, launch_normalize(c->st, nd, (cs_cand_dev*)c->cands.p, cnt, cand_cap));
        uint32_t ncand = 0;
        HIPCHK(c, hipMemcpyAsync(&ncand, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (ncand > cand_cap && attempt == 0) {
            cand_cap = ncand + 1024;
            continue;
        }
        const int64_t nc = std::min<int64_t>(ncand, cand_cap);
        std::vector<cs_cand_dev> hc((size_t)nc);
        if (nc)
            HIPCHK(c, hipMemcpy(hc.data(), c->cands.p, (size_t)nc * sizeof(cs_cand_dev),
                                hipMemcpyDeviceToHost));
        std::sort(hc.begin(), hc.end(), [](const cs_cand_dev& a, const cs_cand_dev& b) {
            return a.t != b.t ? a.t < b.t : a.k < b.k;
        });
        *n_out = nc;
        if (out)
            for (int64_t i = 0; i < std::min<int64_t>(nc, cap); ++i) {
                out[i].t = hc[i].t;
                out[i].k = hc[i].k;
                out[i].nharm = hc[i].nharm;
                out[i].block = hc[i].block;
                out[i].metric = hc[i].metric;
            }
        if (nc > cap) return c->fail(CS_ENOSPC, "candidate buffer too small");
        return CS_OK;
    }
    return c->fail(CS_ENOSPC, "candidate buffer could not be sized");
}

// ---- block triage -----------------------------------------------------------
static_assert(CS_TRIAGE_FEATURES == kTriageFeatures && CS_TRIAGE_PARAMS == kTriageParams,
              "triage sizes of the header and the kernels differ");
int cs_triage_features(cs_ctx* c, double r_lo, double r_hi, float* out, int64_t cap_blocks,
                       int64_t* nblocks) {
    return cs_triage_features_kind(c, 0, r_lo, r_hi, out, cap_blocks, nblocks);
}

int cs_triage_features_kind(cs_ctx* c, int32_t kind, double r_lo, double r_hi, float* out,
                            int64_t cap_blocks, int64_t* nblocks) {
    if (!c || !nblocks) return CS_EINVAL;
    if (kind != 0 && kind != 1) return c->fail(CS_EINVAL, "feature kind must be 0 or 1");
    *nblocks = 0;
    int s = check_spectrum(c);
    if (s) return s;
    if (c->nharms > 64) return c->fail(CS_EINVAL, "triage features need nharms <= 64");
    HIPCHK(c, hipSetDevice(c->device));
    const double lodr = c->hidr / c->nharms;
    const double span = std::floor((r_hi - r_lo) / lodr);
    const int64_t total = span >= 0 ? (int64_t)span + 1 : 0;
    const int64_t nch = (total + c->Nprof - 1) / c->Nprof;
    *nblocks = nch;
    if (nch == 0) return CS_OK;
    if (!out || cap_blocks < nch) return c->fail(CS_ENOSPC, "feature buffer too small");
    uint32_t fm = 0;
    for (int d : c->decims)
        if (d >= 2 && d <= 6) fm |= 1u << d;
    const SpecDev sp = spec_of(c);
    if (kind == 0 &&
        (s = ensure(c, c->pyr, (size_t)triage_pyramid_floats(sp.namps) * sizeof(float))))
        return s;
    if ((s = ensure(c, c->tfeat, (size_t)nch * kTriageFeatures * sizeof(float)))) return s;
    HIPCHK(c, launch_triage_features(c->st, kind, sp.amps, sp.namps, (float*)c->pyr.p, r_lo, lodr,
                                     c->Nprof, total, c->nharms, fm, nch, (float*)c->tfeat.p));
    HIPCHK(c, hipMemcpyAsync(out, c->tfeat.p, (size_t)nch * kTriageFeatures * sizeof(float),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

static int mlp_call(cs_ctx* c, const float* params, const float* norm, const float* X,
                    const float* y, int64_t n, float* logits, float* grad) {
    if (!c || !params || !norm || (n > 0 && !X)) return CS_EINVAL;
    if (n < 0) return c->fail(CS_EINVAL, "n must be >= 0");
    if (n == 0) {
        if (grad) std::memset(grad, 0, sizeof(float) * (kTriageParams + 1));
        return CS_OK;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int s;
    const size_t np = kTriageParams, nx = (size_t)n * kTriageFeatures;
    if ((s = ensure(c, c->mlp_w, (np + 2 * kTriageFeatures) * sizeof(float)))) return s;
    if ((s = ensure(c, c->mlp_x, (nx + 2 * (size_t)n) * sizeof(float)))) return s;
    float* dW = (float*)c->mlp_w.p;
    float* dX = (float*)c->mlp_x.p;
    float* dY = dX + nx;
    float* dZ = dY + n;
    HIPCHK(c, hipMemcpyAsync(dW, params, np * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(dW + np, norm, 2 * kTriageFeatures * sizeof(float),
                             hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(dX, X, nx * sizeof(float), hipMemcpyHostToDevice, c->st));
    float* dP = nullptr;
    float* dG = nullptr;
    if (grad) {
        HIPCHK(c, hipMemcpyAsync(dY, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->st));
        const size_t rows = (size_t)mlp_partial_rows(n);
        if ((s = ensure(c, c->mlp_p, (rows + 1) * (np + 1) * sizeof(float)))) return s;
        dP = (float*)c->mlp_p.p;
        dG = dP + rows * (np + 1);
    }
    HIPCHK(c, launch_mlp(c->st, dW, dW + np, dX, grad ? dY : nullptr, n, logits ? dZ : nullptr, dP,
                         dG));
    if (logits)
        HIPCHK(c, hipMemcpyAsync(logits, dZ, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->st));
    if (grad)
        HIPCHK(c, hipMemcpyAsync(grad, dG, (np + 1) * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

int cs_mlp_forward(cs_ctx* c, const float* params, const float* norm, const float* X, int64_t n,
                   float* logits) {
    if (!logits && n > 0) return CS_EINVAL;
    return mlp_call(c, params, norm, X, nullptr, n, logits, nullptr);
}

int cs_mlp_grad(cs_ctx* c, const float* params, const float* norm, const float* X, const float* y,
                int64_t n, float* grad) {
    if (!grad || (n > 0 && !y)) return CS_EINVAL;
    return mlp_call(c, params, norm, X, y, n, nullptr, grad);
}

int cs_mlp_train(cs_ctx* c, float* params, const float* norm, const float* X, const float* y,
                 int64_t n, int32_t epochs, float lr, float* losses) {
    if (!c || !params || !norm) return CS_EINVAL;
    if (n < 1 || !X || !y) return c->fail(CS_EINVAL, "training needs n >= 1 rows and labels");
    if (epochs < 1 || epochs > 100000) return c->fail(CS_EINVAL, "epochs must be in 1..100000");
    if (!(lr > 0.0f)) return c->fail(CS_EINVAL, "lr must be positive");
    HIPCHK(c, hipSetDevice(c->device));
    int s;
    const size_t np = kTriageParams, nx = (size_t)n * kTriageFeatures;
    // parameters, norm, Adam moments and the loss history
    const size_t nw = np + 2 * kTriageFeatures + 2 * np + (size_t)epochs;
    if ((s = ensure(c, c->mlp_w, nw * sizeof(float)))) return s;
    if ((s = ensure(c, c->mlp_x, (nx + 2 * (size_t)n) * sizeof(float)))) return s;
    const size_t rows = (size_t)mlp_partial_rows(n);
    if ((s = ensure(c, c->mlp_p, (rows + 1) * (np + 1) * sizeof(float)))) return s;
    float* dW = (float*)c->mlp_w.p;
    float* dN = dW + np;
    float* dM1 = dN + 2 * kTriageFeatures;
    float* dM2 = dM1 + np;
    float* dL = dM2 + np;
    float* dX = (float*)c->mlp_x.p;
    float* dY = dX + nx;
    float* dP = (float*)c->mlp_p.p;
    float* dG = dP + rows * (np + 1);
    HIPCHK(c, hipMemcpyAsync(dW, params, np * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(dN, norm, 2 * kTriageFeatures * sizeof(float), hipMemcpyHostToDevice,
                             c->st));
    HIPCHK(c, hipMemcpyAsync(dX, X, nx * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(dY, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemsetAsync(dM1, 0, 2 * np * sizeof(float), c->st));
    // every epoch on the device: forward and backward over all rows (fixed-order
    // reduction), then the Adam step; the host waits once, at the end
    for (int32_t it = 0; it < epochs; ++it) {
        HIPCHK(c, launch_mlp(c->st, dW, dN, dX, dY, n, nullptr, dP, dG));
        HIPCHK(c, launch_adam(c->st, dW, dM1, dM2, dG, (int)np, lr, 0.9f, 0.999f, 1e-8f, it + 1, dL,
                              it));
    }
    HIPCHK(c, hipMemcpyAsync(params, dW, np * sizeof(float), hipMemcpyDeviceToHost, c->st));
    if (losses)
        HIPCHK(c, hipMemcpyAsync(losses, dL, (size_t)epochs * sizeof(float), hipMemcpyDeviceToHost,
                                 c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return CS_OK;
}

static int search_files_impl(cs_ctx* c, int32_t nfiles, const cs_file* files, double threshold,
                             cs_cand* out, int64_t cap, int64_t* counts, int64_t* n_out) {
    *n_out = 0;
    // the sweep is one launch sequence: graph replay (keyed per search) stays off
    struct GraphOff {
        cs_ctx* c;
        bool was;
        ~GraphOff() { c->use_graph = was; }
    } goff{c, c->use_graph};
    c->use_graph = false;
    // the caller's launch shape, restored after the sweep; a length class with
    // a line in the tuning cache runs with the shape cs_autotune stored for it
    struct TuneKeep {
        cs_ctx* c;
        int64_t b;
        int w, o;
        void restore() const {
            c->tune_batch = b;
            c->tune_wg = w;
            c->tune_order = o;
        }
        ~TuneKeep() { restore(); }
    } tkeep{c, c->tune_batch, c->tune_wg, c->tune_order};
    if (nfiles <= 0) return CS_OK;
    if (cap < 0 || (cap > 0 && !out)) return c->fail(CS_EINVAL, "bad candidate buffer");
    int64_t maxn = 0, namps_all = 0;
    std::vector<int64_t> off(nfiles);
    for (int32_t i = 0; i < nfiles; ++i) {
        const cs_file& f = files[i];
        if (!f.amps || f.namps <= 0 || f.N <= 0 || !(f.T > 0))
            return c->fail(CS_EINVAL, "file " + std::to_string(i) + ": invalid spectrum");
        maxn = std::max(maxn, f.namps);
        off[i] = namps_all;
        namps_all += f.namps;
        counts[i] = 0;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int s;
    // every spectrum stays resident for the sweep (3.9 GB for 64 files of
    // 2^22..2^25 samples): an upload never waits for a search to release its slot
    if ((s = ensure(c, c->sweep_amps, (size_t)namps_all * sizeof(float2)))) return s;
    const size_t maxb = (size_t)maxn * sizeof(float2);
    if (c->pin_bytes < maxb) {
        for (void*& p : c->pin) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
        }
        c->pin_bytes = 0;
        for (void*& p : c->pin)
            if (hipHostMalloc(&p, maxb, hipHostMallocDefault) != hipSuccess || !p)
                return c->fail(CS_ENOMEM, "pinned staging allocation failed");
        c->pin_bytes = maxb;
    }
    if (!c->cst) HIPCHK(c, hipStreamCreateWithFlags(&c->cst, hipStreamNonBlocking));
    while ((int32_t)c->sweep_up.size() < nfiles) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->sweep_up.push_back(e);
    }
    for (hipEvent_t& e : c->sweep_ev)
        if (!e) HIPCHK(c, hipEventCreate(&e));
    if ((s = ensure(c, c->sweep_ctr, (size_t)(nfiles + 1) * sizeof(uint32_t)))) return s;

    // files of equal geometry run back to back, so the plan and table caches
    // (keyed on r_lo and the trial count) are rebuilt only between groups;
    // results do not depend on the order
    std::vector<int32_t> order(nfiles);
    for (int32_t i = 0; i < nfiles; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        const cs_file &x = files[a], &y = files[b];
        if (x.N != y.N) return x.N < y.N;
        if (x.r_lo != y.r_lo) return x.r_lo < y.r_lo;
        return x.r_hi < y.r_hi;
    });
    // staging of file k (worker thread): once the upload that last used pinned
    // buffer k & 1 (file k - 2) has completed, a multi-threaded copy into it,
    // then its upload to the file's resident slot on the copy stream
    auto stage = [c, files, &order, &off](int k) -> hipError_t {
        const int32_t fi = order[k];
        const cs_file& f = files[fi];
        const int sl = k & 1;
        hipError_t e = hipSetDevice(c->device);
        if (e == hipSuccess && k >= 2) e = hipEventSynchronize(c->sweep_up[order[k - 2]]);
        if (e != hipSuccess) return e;
        const size_t bytes = (size_t)f.namps * sizeof(float2);
        const unsigned hw = std::thread::hardware_concurrency();
        const int nt = (int)std::max<size_t>(1, std::min<size_t>({(size_t)(hw ? hw : 4), 8,
                                                                   bytes >> 22}));
        std::vector<std::thread> pool;
        const size_t part = (bytes + nt - 1) / nt;
        for (int t = 0; t < nt; ++t) {
            const size_t b0 = std::min(bytes, part * t), b1 = std::min(bytes, b0 + part);
            pool.emplace_back([=] {
                std::memcpy((char*)c->pin[sl] + b0, (const char*)f.amps + b0, b1 - b0);
            });
        }
        for (auto& th : pool) th.join();
        e = hipMemcpyAsync((float2*)c->sweep_amps.p + off[fi], c->pin[sl], bytes,
                           hipMemcpyHostToDevice, c->cst);
        if (e == hipSuccess) e = hipEventRecord(c->sweep_up[fi], c->cst);
        return e;
    };
    uint32_t ccap = (uint32_t)std::max<int64_t>(1 << 18, std::min<int64_t>(cap, (int64_t)1 << 26));
    int rc = CS_OK;
    uint32_t nc_all = 0;
    std::vector<uint32_t> ends(nfiles);
    for (int attempt = 0; attempt < 2; ++attempt) {
        if ((s = ensure(c, c->sweep_cands, (size_t)ccap * sizeof(cs_cand_dev)))) return s;
        SweepSink sk{(cs_cand_dev*)c->sweep_cands.p, (uint32_t*)c->sweep_ctr.p, ccap,
                     (uint32_t*)c->sweep_ctr.p + 1, 0};
        HIPCHK(c, launch_reset_u32(c->st, sk.count, 1));
        HIPCHK(c, hipEventRecord(c->sweep_ev[0], c->st));
        std::future<hipError_t> next;
        if (attempt == 0) next = std::async(std::launch::async, stage, 0);
        for (int32_t k = 0; k < nfiles && rc == CS_OK; ++k) {
            if (next.valid()) {
                const hipError_t e = next.get();
                if (e != hipSuccess) {
                    rc = c->fail(CS_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
                    break;
                }
            }
            if (attempt == 0 && k + 1 < nfiles) next = std::async(std::launch::async, stage, k + 1);
            const int32_t fi = order[k];
            const cs_file& f = files[fi];
            hipError_t e = hipStreamWaitEvent(c->st, c->sweep_up[fi], 0);
            if (e != hipSuccess) {
                rc = c->fail(CS_EDEVICE, std::string("upload wait: ") + hipGetErrorString(e));
                break;
            }
            c->amps_view = (const float2*)c->sweep_amps.p + off[fi];
            c->namps = f.namps;
            c->N = f.N;
            c->T = f.T;
            c->spl_valid = false;
            if (!c->tune_cache.empty() &&
                (k == 0 || files[order[k - 1]].N != f.N || files[order[k - 1]].r_lo != f.r_lo ||
                 files[order[k - 1]].r_hi != f.r_hi)) {
                int64_t b = 0;
                int w = 0, o = 0;
                tkeep.restore();
                if (tune_lookup(c, c->tune_cache, f.r_lo, f.r_hi, b, w, o)) {
                    c->tune_batch = b;
                    c->tune_wg = w;
                    c->tune_order = o;
                }
            }
            sk.file = fi;
            int64_t n = 0;
            const int st = search_impl(c, f.r_lo, f.r_hi, threshold, nullptr, 0, nullptr, 0, &n,
                                       nullptr, nullptr, &sk);
            if (st != CS_OK) rc = st;
        }
        if (next.valid()) (void)next.get();
        HIPCHK(c, hipEventRecord(c->sweep_ev[1], c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (rc != CS_OK) break;
        HIPCHK(c, hipMemcpy(&nc_all, sk.count, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (nc_all <= ccap) {
            HIPCHK(c, hipMemcpy(ends.data(), sk.ends, (size_t)nfiles * sizeof(uint32_t),
                                hipMemcpyDeviceToHost));
            break;
        }
        ccap = nc_all + 1024;   // the spectra are still resident: search again
    }
    (void)hipStreamSynchronize(c->cst);
    c->amps_view = nullptr;
    c->namps = 0;
    c->spl_valid = false;
    if (rc != CS_OK) return rc;
    for (double& v : c->timings) v = 0.0;
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->sweep_ev[0], c->sweep_ev[1]));
    c->timings[6] = ms;
    std::vector<cs_cand_dev> hc((size_t)nc_all);
    if (nc_all)
        HIPCHK(c, hipMemcpy(hc.data(), c->sweep_cands.p, (size_t)nc_all * sizeof(cs_cand_dev),
                            hipMemcpyDeviceToHost));
    // file order[k] holds records ends[order[k - 1]] .. ends[order[k]] - 1
    std::vector<std::pair<uint32_t, uint32_t>> rng(nfiles);
    uint32_t prev = 0;
    for (int32_t k = 0; k < nfiles; ++k) {
        rng[order[k]] = {prev, ends[order[k]]};
        prev = ends[order[k]];
    }
    int64_t total = 0;
    for (int32_t i = 0; i < nfiles; ++i) {
        auto b0 = hc.begin() + rng[i].first, b1 = hc.begin() + rng[i].second;
        std::sort(b0, b1, [](const cs_cand_dev& a, const cs_cand_dev& b) {
            return a.t != b.t ? a.t < b.t : a.k < b.k;
        });
        counts[i] = (int64_t)(rng[i].second - rng[i].first);
        for (auto it = b0; it != b1; ++it) {
            if (total < cap) {
                out[total].t = it->t;
                out[total].k = it->k;
                out[total].nharm = it->nharm;
                out[total].block = it->block;
                out[total].metric = it->metric;
            }
            ++total;
        }
    }
    *n_out = total;
    return total > cap ? CS_ENOSPC : CS_OK;
}

int cs_search_files(cs_ctx* c, int32_t nfiles, const cs_file* files, double threshold,
                    cs_cand* out, int64_t cap, int64_t* counts, int64_t* n_out) {
    if (!c || !n_out || (nfiles > 0 && (!files || !counts))) return CS_EINVAL;
    try {
        return search_files_impl(c, nfiles, files, threshold, out, cap, counts, n_out);
    } catch (const std::exception& e) {
        c->amps_view = nullptr;
        c->namps = 0;
        return c->fail(CS_ENOMEM, std::string("multi-file search: ") + e.what());
    }
}

}  // extern "C"
