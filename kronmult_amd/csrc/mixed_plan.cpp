// mixed_plan.cpp -- host side of the mixed-(d, n) persistent launch: the plan
// (Y-target groups, overlap proof, cost-sorted work items), the autotuner and
// its on-disk cache, and hipGraph capture of repeated applies.  Host code
// only; the kernel is in kron_mixed.hip.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "kron_mixed.hpp"

namespace kron {

namespace {

// A group longer than this is split across work items (the class then
// updates Y atomically), so one hot Y target cannot serialise the launch.
constexpr int kMaxGroup = 64;
// Scratch of the classes that run as separate launches is capped here.
constexpr size_t kLaunchScratchBytes = size_t(1) << 30;
// Whole-GPU HBM rate assumed by the untuned cost model (bytes per microsecond).
constexpr double kModelBytesPerUs = 4.0e6;

struct ClassDesc {
    int d = 0, n = 0, lda = 0, batch = 0;
    const void *pA = nullptr;
    const void *pX = nullptr;
    void *pY = nullptr;
};

struct Tuned {
    int variant;
    double us;  // whole-GPU microseconds per entry
};

std::mutex g_tune_mu;
std::map<std::tuple<std::string, int, int, int>, Tuned> g_tuned;  // (arch, elem, d, n)

std::string device_arch()
{
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
        return "unknown";
    std::string a = p.gcnArchName;
    const size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}

// Cache file: one "arch elem d n variant us_per_entry" line per class.
void load_cache(const char *path)
{
    if (!path) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    char arch[64];
    int elem, d, n, v;
    double us;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    while (std::fscanf(f, "%63s %d %d %d %d %lf", arch, &elem, &d, &n, &v, &us) == 6) {
        if (v >= 0 && v < kMixVariants && us > 0)
            g_tuned[std::make_tuple(std::string(arch), elem, d, n)] = Tuned{v, us};
    }
    std::fclose(f);
}

void append_cache(const char *path, const std::string &arch, int elem, int d, int n,
                  const Tuned &t)
{
    if (!path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d %d %.6g\n", arch.c_str(), elem, d, n, t.variant, t.us);
    std::fclose(f);
}

bool lookup(const std::string &arch, int elem, int d, int n, Tuned &out)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(arch, elem, d, n));
    if (it == g_tuned.end()) return false;
    out = it->second;
    return true;
}

long long ipow_ll(int n, int d)
{
    long long r = 1;
    for (int k = 0; k < d; ++k) r *= n;
    return r;
}

template <typename T>
double model_cost_us(int d, int n, const MixPathInfo &info)
{
    const double E = (double)ipow_ll(n, d);
    double bytes = sizeof(T) * (3.0 * E + (double)d * n * n) + 8.0 * (d + 2);
    if (info.big) bytes += 2.0 * sizeof(T) * E;  // the scratch round trip
    // slab classes read X n times, all but once mostly from the caches
    if (info.slab) bytes += 0.25 * (n - 1) * sizeof(T) * E;
    return bytes / kModelBytesPerUs;
}

hipError_t upload(void **dst, const void *src, size_t bytes)
{
    *dst = nullptr;
    if (bytes == 0) return hipSuccess;
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
}

}  // namespace

struct MixedPlanCore {
    int elem = 0;
    int device = 0;
    struct Cls {
        ClassDesc desc;
        int kind = KRON_MIXED_KIND_LAUNCH;
        int variant = 0;
        int atomic = 1;
        double cost = 0;
    };
    std::vector<Cls> cls;
    MixedLaunch L[2];  // pass-1 tiles + part of the rest, pass-2 tiles + the rest
    std::vector<void *> allocs;
    // classes outside the compiled set
    std::vector<KronArgs<double>> launch64;
    std::vector<KronArgs<float>> launch32;
    // graph of repeated applies
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~MixedPlanCore()
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        for (void *p : allocs)
            if (p) (void)hipFree(p);
        (void)hipSetDevice(cur);
    }
    void *keep(void *p)
    {
        allocs.push_back(p);
        return p;
    }
};

template <typename T>
static std::vector<KronArgs<T>> &launch_list(MixedPlanCore &p);
template <>
std::vector<KronArgs<double>> &launch_list<double>(MixedPlanCore &p) { return p.launch64; }
template <>
std::vector<KronArgs<float>> &launch_list<float>(MixedPlanCore &p) { return p.launch32; }

template <typename T>
static kron_status plan_execute_t(MixedPlanCore &p, hipStream_t s)
{
    kron_status st = KRON_OK;
    for (const MixedLaunch &L : p.L)
        if ((st = mixed_launch<T>(L, s)) != KRON_OK) return st;
    for (KronArgs<T> a : launch_list<T>(p)) {
        a.stream = s;
        st = launch_kron(a);
        if (st != KRON_OK) return st;
    }
    return KRON_OK;
}

kron_status mixed_execute(MixedPlanCore *p, hipStream_t s)
{
    if (!p) return KRON_EINVAL;
    return p->elem == 8 ? plan_execute_t<double>(*p, s) : plan_execute_t<float>(*p, s);
}

// Build a plan.  forced: -1 = tuned / default mapping, else that variant
// where it exists.  tune: time the mappings of uncached classes.
template <typename T>
static kron_status build_plan(MixedPlanCore *&out, const std::vector<ClassDesc> &classes,
                              const char *cache, int flags, hipStream_t stream);

// Time every mapping of class (d, n) on synthetic data and keep the fastest.
template <typename T>
static kron_status tune_class(int d, int n, hipStream_t stream, Tuned &best)
{
    const long long E = ipow_ll(n, d);
    long long nent = (256LL << 20) / (2 * E * (long long)sizeof(T));
    nent = std::max(1024LL, std::min(16384LL, nent));
    const size_t xbytes = (size_t)nent * E * sizeof(T);
    const size_t abytes = (size_t)nent * d * n * n * sizeof(T);
    void *X = nullptr, *Y = nullptr, *A = nullptr, *ptrs = nullptr;
    kron_status st = KRON_OK;
    std::vector<const T *> hA((size_t)nent * d), hX(nent);
    std::vector<T *> hY(nent);
    auto fail = [&](hipError_t e) {
        set_last_hip_error((int)e);
        st = KRON_EHIP;
    };
    hipError_t e;
    if ((e = hipMalloc(&X, xbytes)) != hipSuccess || (e = hipMalloc(&Y, xbytes)) != hipSuccess ||
        (e = hipMalloc(&A, abytes)) != hipSuccess ||
        (e = hipMalloc(&ptrs, sizeof(void *) * (size_t)nent * (d + 2))) != hipSuccess) {
        fail(e);
    } else {
        (void)hipMemsetAsync(X, 0, xbytes, stream);
        (void)hipMemsetAsync(Y, 0, xbytes, stream);
        (void)hipMemsetAsync(A, 0, abytes, stream);
        for (long long b = 0; b < nent; ++b) {
            for (int k = 0; k < d; ++k)
                hA[b * d + k] = static_cast<const T *>(A) + (b * d + k) * n * n;
            hX[b] = static_cast<const T *>(X) + b * E;
            hY[b] = static_cast<T *>(Y) + b * E;
        }
        void **pp = static_cast<void **>(ptrs);
        if ((e = hipMemcpy(pp, hA.data(), sizeof(void *) * nent * d, hipMemcpyHostToDevice)) !=
                hipSuccess ||
            (e = hipMemcpy(pp + nent * d, hX.data(), sizeof(void *) * nent,
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(pp + nent * (d + 1), hY.data(), sizeof(void *) * nent,
                           hipMemcpyHostToDevice)) != hipSuccess)
            fail(e);
    }
    best = Tuned{-1, 0};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st == KRON_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = KRON_EHIP;
    for (int v = 0; v < kMixVariants && st == KRON_OK; ++v) {
        MixPathInfo info;
        if (!mixed_path_info<T>(d, n, v, info)) continue;
        // two-pass mappings are not candidates: with one of them an apply
        // takes two launches, and the slab mappings beat them
        if (info.big) continue;
        void **pp = static_cast<void **>(ptrs);
        std::vector<ClassDesc> one(1);
        one[0] = ClassDesc{d, n, n, (int)nent, pp, pp + nent * d, pp + nent * (d + 1)};
        MixedPlanCore *sub = nullptr;
        const int fl = KRON_MIXED_FORCE_V0 << v;
        st = build_plan<T>(sub, one, nullptr, fl, stream);
        if (st != KRON_OK) break;
        const int reps = 3;
        float ms = 0;
        st = mixed_execute(sub, stream);  // warm-up
        if (st == KRON_OK) {
            (void)hipEventRecord(e0, stream);
            for (int r = 0; r < reps && st == KRON_OK; ++r) st = mixed_execute(sub, stream);
            (void)hipEventRecord(e1, stream);
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) !=
                                                             hipSuccess)
                st = KRON_EHIP;
        }
        delete sub;
        if (st != KRON_OK) break;
        const double us = 1000.0 * ms / reps / (double)nent;
        if (best.variant < 0 || us < best.us) best = Tuned{v, us};
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(stream);
    for (void *q : {X, Y, A, ptrs})
        if (q) (void)hipFree(q);
    if (st == KRON_OK && best.variant < 0) st = KRON_ENOTSUP;
    return st;
}

template <typename T>
static kron_status build_plan(MixedPlanCore *&out, const std::vector<ClassDesc> &classes,
                              const char *cache, int flags, hipStream_t stream)
{
    out = nullptr;
    int forced = -1;
    for (int v = kMixVariants - 1; v >= 0; --v)
        if (flags & (KRON_MIXED_FORCE_V0 << v)) forced = v;
    const std::string arch = device_arch();
    const int nclass = (int)classes.size();
    auto *p = new MixedPlanCore();
    p->elem = (int)sizeof(T);
    (void)hipGetDevice(&p->device);
    p->cls.resize(nclass);
    auto bail = [&](kron_status st) {
        delete p;
        return st;
    };

    // 1. each class's path and mapping
    if (cache) load_cache(cache);
    size_t lds = 16;
    for (int c = 0; c < nclass; ++c) {
        const ClassDesc &cd = classes[c];
        MixedPlanCore::Cls &C = p->cls[c];
        C.desc = cd;
        MixPathInfo info;
        int v = -1;
        Tuned t{0, 0};
        const bool cached = lookup(arch, (int)sizeof(T), cd.d, cd.n, t);
        if (forced >= 0 && mixed_path_info<T>(cd.d, cd.n, forced, info)) {
            v = forced;
        } else if (forced < 0 && cached && mixed_path_info<T>(cd.d, cd.n, t.variant, info)) {
            v = t.variant;
        } else if (forced < 0 && (flags & KRON_MIXED_TUNE) && cd.batch > 0 &&
                   mixed_path_info<T>(cd.d, cd.n, 0, info)) {
            kron_status st = tune_class<T>(cd.d, cd.n, stream, t);
            if (st != KRON_OK) return bail(st);
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_tuned[std::make_tuple(arch, (int)sizeof(T), cd.d, cd.n)] = t;
            }
            append_cache(cache, arch, (int)sizeof(T), cd.d, cd.n, t);
            mixed_path_info<T>(cd.d, cd.n, t.variant, info);
            v = t.variant;
        } else if (mixed_path_info<T>(cd.d, cd.n, 0, info)) {
            v = 0;
            t.us = 0;
        }
        if (v < 0) {
            C.kind = KRON_MIXED_KIND_LAUNCH;
            continue;
        }
        C.variant = v;
        C.kind = info.slab ? KRON_MIXED_KIND_SLAB
                 : info.big ? KRON_MIXED_KIND_TWOPASS
                 : info.packed ? KRON_MIXED_KIND_PACKED : KRON_MIXED_KIND_ONCHIP;
        C.cost = (t.us > 0 && t.variant == v) ? t.us : model_cost_us<T>(cd.d, cd.n, info);
        C.atomic = (info.packed || (flags & KRON_MIXED_ATOMIC)) ? 1 : 0;
        lds = std::max(lds, info.lds);
    }

    // 2. Y targets on the host: group each class's entries by target and prove
    //    distinct targets disjoint (else the classes involved go atomic)
    std::vector<std::vector<unsigned long long>> ykeys(nclass);
    std::vector<std::vector<int>> sorted(nclass);
    struct Iv {
        unsigned long long lo, hi;
        int cls;
    };
    std::vector<Iv> ivs;
    for (int c = 0; c < nclass; ++c) {
        const MixedPlanCore::Cls &C = p->cls[c];
        if (C.kind == KRON_MIXED_KIND_LAUNCH || C.desc.batch == 0) continue;
        ykeys[c].resize(C.desc.batch);
        hipError_t e = hipMemcpy(ykeys[c].data(), C.desc.pY,
                                 sizeof(unsigned long long) * C.desc.batch, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            set_last_hip_error((int)e);
            return bail(KRON_EHIP);
        }
        const unsigned long long bytes =
            (unsigned long long)ipow_ll(C.desc.n, C.desc.d) * sizeof(T);
        std::vector<int> &ord = sorted[c];
        ord.resize(C.desc.batch);
        std::iota(ord.begin(), ord.end(), 0);
        const auto &keys = ykeys[c];
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return keys[a] < keys[b]; });
        for (size_t i = 0; i < ord.size(); ++i)
            if (i == 0 || keys[ord[i]] != keys[ord[i - 1]])
                ivs.push_back(Iv{keys[ord[i]], keys[ord[i]] + bytes, c});
    }
    std::sort(ivs.begin(), ivs.end(), [](const Iv &a, const Iv &b) {
        return a.lo != b.lo ? a.lo < b.lo : a.cls < b.cls;
    });
    // sorted by start, every later interval that starts before interval i
    // ends overlaps it
    for (size_t i = 0; i < ivs.size(); ++i)
        for (size_t j = i + 1; j < ivs.size() && ivs[j].lo < ivs[i].hi; ++j) {
            p->cls[ivs[i].cls].atomic = 1;
            p->cls[ivs[j].cls].atomic = 1;
        }

    // 3. work items.  Launch 0 holds the pass-1 tiles of the two-pass
    //    classes, launch 1 their pass-2 tiles; the whole-entry items go to
    //    whichever launch is lighter (largest first), so both stay full.
    std::vector<MixedItem> items;
    std::vector<double> icost;
    std::vector<int> ilaunch;
    std::vector<int> order;       // class-local entry index of each record
    std::vector<int> order_cls;   // class of each record
    long long wtotal = 0;  // scratch elements of the two-pass classes
    std::vector<long long> wofs(nclass, 0);
    double load[2] = {0, 0};
    std::vector<int> whole;  // indices of whole-entry items
    for (int c = 0; c < nclass; ++c) {
        MixedPlanCore::Cls &C = p->cls[c];
        if (C.kind == KRON_MIXED_KIND_LAUNCH || C.desc.batch == 0) continue;
        const std::vector<int> &ord = sorted[c];
        const auto &keys = ykeys[c];
        const int base = (int)order.size();
        order.insert(order.end(), ord.begin(), ord.end());
        order_cls.insert(order_cls.end(), ord.size(), c);
        MixPathInfo info;
        mixed_path_info<T>(C.desc.d, C.desc.n, C.variant, info);
        auto add = [&](MixedItem it, double cost, int launch) {
            items.push_back(it);
            icost.push_back(cost);
            ilaunch.push_back(launch);
            if (launch < 0)
                whole.push_back((int)items.size() - 1);
            else
                load[launch] += cost;
        };
        if (C.kind == KRON_MIXED_KIND_PACKED) {
            for (int i = 0; i < C.desc.batch; i += info.per_item) {
                const int cnt = std::min(info.per_item, C.desc.batch - i);
                add(MixedItem{c, base + i, cnt, -1}, C.cost * cnt, -1);
            }
            continue;
        }
        // a group is one item (plain Y updates) unless the class is atomic:
        // then single entries, or chunks of kMaxGroup when groups are long
        int longest = 0;
        for (int i = 0; i < C.desc.batch;) {
            int j = i + 1;
            while (j < C.desc.batch && keys[ord[j]] == keys[ord[i]]) ++j;
            longest = std::max(longest, j - i);
            i = j;
        }
        if (longest > kMaxGroup) C.atomic = 1;
        const int cap = !C.atomic ? C.desc.batch : longest > kMaxGroup ? kMaxGroup : 1;
        // a two-pass entry moves X in and W out in pass 1 (2 vectors), W and
        // Y in and Y out in pass 2 (3 vectors): its cost splits 2 : 3
        if (C.kind == KRON_MIXED_KIND_TWOPASS) {
            // record r of the class (r >= base) uses scratch wofs + r * n^d
            wofs[c] = wtotal - (long long)base * ipow_ll(C.desc.n, C.desc.d);
            wtotal += (long long)C.desc.batch * ipow_ll(C.desc.n, C.desc.d);
            for (int i = 0; i < C.desc.batch; ++i)
                for (int t = 0; t < info.tiles1; ++t)
                    add(MixedItem{c, base + i, 1, t}, 0.4 * C.cost / info.tiles1, 0);
        }
        for (int i = 0; i < C.desc.batch;) {
            int j = i + 1;
            while (j < C.desc.batch && j - i < cap && keys[ord[j]] == keys[ord[i]]) ++j;
            if (C.kind == KRON_MIXED_KIND_TWOPASS) {
                for (int t = 0; t < info.tiles2; ++t)
                    add(MixedItem{c, base + i, j - i, -2 - t},
                        0.6 * C.cost * (j - i) / info.tiles2, 1);
            } else if (C.kind == KRON_MIXED_KIND_SLAB) {
                // item t holds slabs t*S .. t*S+S-1 (S = info.slabs); a
                // group's items are neighbours in the queue (equal costs,
                // stable sort)
                const int S = info.slabs;
                for (int t = 0; t < info.tiles1; ++t)
                    add(MixedItem{c, base + i, j - i, t * S}, C.cost * (j - i) / info.tiles1, -1);
            } else {
                add(MixedItem{c, base + i, j - i, -1}, C.cost * (j - i), -1);
            }
            i = j;
        }
    }
    std::stable_sort(whole.begin(), whole.end(), [&](int a, int b) { return icost[a] > icost[b]; });
    const bool two = wtotal > 0;
    for (int w : whole) {
        const int l = (two && load[1] < load[0]) ? 1 : 0;
        ilaunch[w] = l;
        load[l] += icost[w];
    }
    // each launch's items on one queue per XCD: the slab items of one group
    // form a bundle that stays on one queue (they all read the group's X, so
    // on one XCD they share its L2); bundles go to the least loaded queue,
    // largest first, and each queue runs its items largest first (a bundle's
    // items stay neighbours: equal costs, stable sort)
    std::vector<MixedItem> sitems;
    int nA = 0;
    MixQueues qs[2] = {};
    for (int l = 0; l < 2; ++l) {
        std::vector<std::vector<int>> bundles;
        std::map<std::pair<int, int>, int> slab_bundle;
        for (size_t i = 0; i < items.size(); ++i) {
            if (ilaunch[i] != l) continue;
            const MixedItem &it = items[i];
            if (p->cls[it.cls].kind == KRON_MIXED_KIND_SLAB) {
                auto ins = slab_bundle.emplace(std::make_pair(it.cls, it.first),
                                               (int)bundles.size());
                if (ins.second) bundles.emplace_back();
                bundles[ins.first->second].push_back((int)i);
            } else {
                bundles.push_back({(int)i});
            }
        }
        std::vector<double> bcost(bundles.size(), 0.0);
        std::vector<int> bord(bundles.size());
        for (size_t b = 0; b < bundles.size(); ++b) {
            bord[b] = (int)b;
            for (int i : bundles[b]) bcost[b] += icost[i];
        }
        std::stable_sort(bord.begin(), bord.end(),
                         [&](int a, int b) { return bcost[a] > bcost[b]; });
        std::vector<std::vector<int>> queue(kMixXcds);
        double qload[kMixXcds] = {};
        for (int b : bord) {
            int q = 0;
            for (int k = 1; k < kMixXcds; ++k)
                if (qload[k] < qload[q]) q = k;
            qload[q] += bcost[b];
            queue[q].insert(queue[q].end(), bundles[b].begin(), bundles[b].end());
        }
        const int base = (int)sitems.size();
        for (int q = 0; q < kMixXcds; ++q) {
            std::stable_sort(queue[q].begin(), queue[q].end(),
                             [&](int a, int b) { return icost[a] > icost[b]; });
            qs[l].off[q] = (int)sitems.size() - base;
            for (int i : queue[q]) sitems.push_back(items[i]);
        }
        qs[l].off[kMixXcds] = (int)sitems.size() - base;
        if (l == 0) nA = (int)sitems.size();
    }

    // 4. device tables
    std::vector<MixedClass<T>> dcls(nclass);
    for (int c = 0; c < nclass; ++c) {
        const MixedPlanCore::Cls &C = p->cls[c];
        MixedClass<T> &m = dcls[c];
        m.pA = static_cast<const T *const *>(C.desc.pA);
        m.pX = static_cast<const T *const *>(C.desc.pX);
        m.pY = static_cast<T *const *>(C.desc.pY);
        m.d = C.desc.d;
        m.n = C.desc.n;
        m.lda = C.desc.lda;
        m.path = C.kind == KRON_MIXED_KIND_LAUNCH ? -1
                                                  : mixed_path_code(C.desc.d, C.desc.n, C.variant);
        m.atomic_y = C.atomic;
        m.pad = 0;
        m.wofs = wofs[c];
    }
    hipError_t e = hipSuccess;
    void *dc = nullptr, *di = nullptr, *dord = nullptr, *dq = nullptr, *dscr = nullptr;
    auto put = [&](void **dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = upload(dst, src, bytes);
        if (*dst) p->keep(*dst);
    };
    put(&dc, dcls.data(), sizeof(MixedClass<T>) * dcls.size());
    put(&di, sitems.data(), sizeof(MixedItem) * sitems.size());
    // the entries' pointers, gathered once from the caller's pointer arrays
    std::vector<MixedEntry> recs(order.size());
    for (int c = 0; c < nclass; ++c) {
        const MixedPlanCore::Cls &C = p->cls[c];
        if (C.kind == KRON_MIXED_KIND_LAUNCH || C.desc.batch == 0) continue;
        std::vector<const void *> hA((size_t)C.desc.batch * C.desc.d), hX(C.desc.batch);
        if ((e = hipMemcpy(hA.data(), C.desc.pA, sizeof(void *) * hA.size(),
                           hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(hX.data(), C.desc.pX, sizeof(void *) * hX.size(),
                           hipMemcpyDeviceToHost)) != hipSuccess) {
            set_last_hip_error((int)e);
            return bail(KRON_EHIP);
        }
        for (size_t r = 0; r < order.size(); ++r) {
            if (order_cls[r] != c) continue;
            const int b = order[r];
            MixedEntry &m = recs[r];
            for (int k = 0; k < kMaxDim; ++k)
                m.A[k] = k < C.desc.d ? hA[(size_t)b * C.desc.d + k] : nullptr;
            m.X = hX[b];
            m.Y = reinterpret_cast<void *>(ykeys[c][b]);
        }
    }
    put(&dord, recs.data(), sizeof(MixedEntry) * recs.size());
    if (e == hipSuccess &&
        (e = hipMalloc(&dq, 2 * kMixQueueWords * sizeof(unsigned))) == hipSuccess) {
        p->keep(dq);
        e = hipMemset(dq, 0, 2 * kMixQueueWords * sizeof(unsigned));
    }
    if (e == hipSuccess && wtotal > 0 && (e = hipMalloc(&dscr, wtotal * sizeof(T))) == hipSuccess)
        p->keep(dscr);
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return bail(KRON_EHIP);
    }
    const long long resident = persistent_blocks(mixed_kernel<T>(), lds);
    for (int l = 0; l < 2; ++l) {
        MixedLaunch &L = p->L[l];
        L.classes = dc;
        L.items = static_cast<const MixedItem *>(di) + (l == 0 ? 0 : nA);
        L.recs = static_cast<const MixedEntry *>(dord);
        L.nitems = l == 0 ? nA : (int)sitems.size() - nA;
        L.queue = static_cast<unsigned *>(dq) + kMixQueueWords * l;
        L.q = qs[l];
        L.lds = lds;
        long long grid = resident > 0 ? resident : 1024;
        grid = std::max(1LL, std::min(grid, (long long)L.nitems));
        L.grid = (int)grid;
        L.scratch = dscr;
    }

    // 5. classes outside the compiled set: their own launches, plan-owned scratch
    for (int c = 0; c < nclass; ++c) {
        const MixedPlanCore::Cls &C = p->cls[c];
        if (C.kind != KRON_MIXED_KIND_LAUNCH || C.desc.batch == 0) continue;
        KronArgs<T> a;
        a.d = C.desc.d;
        for (int k = 0; k < kMaxDim; ++k) a.m[k] = a.n[k] = (k < a.d) ? C.desc.n : 1;
        a.lda = C.desc.lda;
        a.layout = Layout::Ptr;
        a.pA = static_cast<const T *const *>(C.desc.pA);
        a.pX = static_cast<const T *const *>(C.desc.pX);
        a.pY = static_cast<T *const *>(C.desc.pY);
        a.y_shared = true;
        a.batch = C.desc.batch;
        const size_t per = kron_scratch_elems(a);
        if (per == (size_t)-1) return bail(KRON_ENOTSUP);
        if (per > 0) {
            long long slots = (long long)(kLaunchScratchBytes / (per * sizeof(T)));
            slots = std::max(1LL, std::min(slots, (long long)a.batch));
            void *w = nullptr;
            if ((e = hipMalloc(&w, (size_t)slots * per * sizeof(T))) != hipSuccess) {
                set_last_hip_error((int)e);
                return bail(KRON_EHIP);
            }
            p->keep(w);
            a.Wbase = static_cast<T *>(w);
            a.slot_stride = per;
            a.nslots = (int)slots;
            a.scratch = Scratch::Linear;
        }
        launch_list<T>(*p).push_back(a);
    }
    out = p;
    return KRON_OK;
}

template <typename C, typename T>
static kron_status plan_create_t(MixedPlanCore *&out, int nclass, const C *classes,
                                 const char *cache, int flags, hipStream_t stream)
{
    out = nullptr;
    if (nclass < 0 || (nclass > 0 && !classes)) return KRON_EINVAL;
    std::vector<ClassDesc> v(nclass);
    for (int c = 0; c < nclass; ++c) {
        const C &s = classes[c];
        if (s.d < 1 || s.d > kMaxDim || s.n < 1 || s.batchCount < 0) return KRON_EINVAL;
        long long E = 1;
        for (int k = 0; k < s.d; ++k) {
            E *= s.n;
            if (E > 0x7fffffffLL) return KRON_EINVAL;
        }
        const int lda = s.lda == 0 ? s.n : s.lda;
        if (lda < s.n) return KRON_EINVAL;
        if (s.batchCount > 0 && (!s.Aarray || !s.pX || !s.pY)) return KRON_EINVAL;
        // the pointer arrays themselves are only read
        v[c] = ClassDesc{s.d, s.n, lda, s.batchCount, s.Aarray, s.pX,
                         const_cast<void *>(static_cast<const void *>(s.pY))};
    }
    return build_plan<T>(out, v, cache, flags, stream);
}

kron_status mixed_create_f64(MixedPlanCore *&out, int nclass, const kron_mixed_class_f64 *c,
                             const char *cache, int flags, hipStream_t s)
{
    return plan_create_t<kron_mixed_class_f64, double>(out, nclass, c, cache, flags, s);
}

kron_status mixed_create_f32(MixedPlanCore *&out, int nclass, const kron_mixed_class_f32 *c,
                             const char *cache, int flags, hipStream_t s)
{
    return plan_create_t<kron_mixed_class_f32, float>(out, nclass, c, cache, flags, s);
}

kron_status mixed_capture(MixedPlanCore *p, int applies)
{
    if (!p || applies < 1) return KRON_EINVAL;
    hipError_t e;
    if (!p->cap_stream &&
        (e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking)) != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    if (p->exec) {
        (void)hipGraphExecDestroy(p->exec);
        p->exec = nullptr;
    }
    if (p->graph) {
        (void)hipGraphDestroy(p->graph);
        p->graph = nullptr;
    }
    if ((e = hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal)) !=
        hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    kron_status st = KRON_OK;
    for (int i = 0; i < applies && st == KRON_OK; ++i) st = mixed_execute(p, p->cap_stream);
    e = hipStreamEndCapture(p->cap_stream, &p->graph);
    if (st != KRON_OK) return st;
    if (e != hipSuccess || (e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0)) !=
                               hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

kron_status mixed_replay(MixedPlanCore *p, hipStream_t s)
{
    if (!p || !p->exec) return KRON_EINVAL;
    hipError_t e = hipGraphLaunch(p->exec, s);
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

kron_status mixed_class_info(MixedPlanCore *p, int cls, int *kind, int *variant, int *atomic,
                             double *cost)
{
    if (!p || cls < 0 || cls >= (int)p->cls.size()) return KRON_EINVAL;
    const MixedPlanCore::Cls &C = p->cls[cls];
    if (kind) *kind = C.kind;
    if (variant) *variant = C.variant;
    if (atomic) *atomic = C.atomic;
    if (cost) *cost = C.cost;
    return KRON_OK;
}

kron_status mixed_launch_info(MixedPlanCore *p, int *items, int *grid, int *lds, int *launches)
{
    if (!p) return KRON_EINVAL;
    if (launches)
        *launches = (p->L[0].nitems > 0) + (p->L[1].nitems > 0) +
                    (int)(p->elem == 8 ? p->launch64.size() : p->launch32.size());
    if (items) *items = p->L[0].nitems + p->L[1].nitems;
    if (grid) *grid = std::max(p->L[0].grid, p->L[1].grid);
    if (lds) *lds = (int)p->L[0].lds;
    return KRON_OK;
}

void mixed_destroy(MixedPlanCore *p) { delete p; }

}  // namespace kron
