/* Host planning with no HIP in it (tmc_plan.h). */
#include "tmc_plan.h"

#include <algorithm>
#include <cmath>
#include <iterator>

/* Parameter checks.  The reference has none (params.h:14 only warns that
 * MU_A must be non-zero: with mu_a = 0 the weight never reaches the cutoff
 * and photon() never returns). */
int tmc::check_params(const tmc_params* p)
{
    if (!p)
        return TMC_E_INVALID;
    if (!std::isfinite(p->mu_a) || !std::isfinite(p->mu_s) || !std::isfinite(p->microns_per_shell))
        return TMC_E_INVALID;
    if (!(p->mu_a > 0.0f) || !(p->mu_s >= 0.0f) || !(p->microns_per_shell > 0.0f))
        return TMC_E_INVALID;
    if (p->shells < 1 || p->shells > TMC_MAX_SHELLS)
        return TMC_E_INVALID;
    /* binary32 albedo must stay below 1, or the weight never decays */
    const volatile float sum = p->mu_s + p->mu_a;
    const volatile float albedo = p->mu_s / sum;
    if (!(albedo < 1.0f))
        return TMC_E_INVALID;
    const double spm = 1e4 / (double)p->microns_per_shell / (double)sum;
    if (!std::isfinite((float)spm) || !((float)spm > 0.0f))
        return TMC_E_INVALID;
    return TMC_OK;
}

/* photon.c:8: albedo evaluated in float. */
float tmc::derived_albedo(const tmc_params* p)
{
    const volatile float sum = p->mu_s + p->mu_a;
    return p->mu_s / sum;
}

/* photon.c:9: 1e4 is a double literal, so the quotient is formed in double
 * from the float sum mu_a + mu_s, then stored as float. */
float tmc::derived_spm(const tmc_params* p)
{
    const volatile float sum = p->mu_a + p->mu_s;
    return (float)(1e4 / (double)p->microns_per_shell / (double)sum);
}

/* The weight after k absorptions is a deterministic binary32 sequence
 * (photon.c:32, 45-49): it falls below 0.001 after k1 steps, then after k2
 * steps of each later cycle (the first of them; later cycles can differ by
 * a step).  False if the first cycle is longer than `limit` steps. */
bool tmc::weight_cadence(const tmc_params* p, uint32_t limit, uint32_t* k1_out, uint32_t* k2_out)
{
    const float a = derived_albedo(p);
    float w = 1.0f;
    uint32_t k1 = 0;
    while (w >= 0.001f && k1 < limit) {
        w = w * a;
        ++k1;
    }
    if (!(w < 0.001f))
        return false;
    /* later cycles start from the cutoff weight / 0.1 */
    float v = w / 0.1f;
    uint32_t k2 = 0;
    while (v >= 0.001f && k2 < limit) {
        v = v * a;
        ++k2;
    }
    *k1_out = k1;
    *k2_out = k2 ? k2 : 1;
    return true;
}

/* Mean absorptions k1 + k2 p / (1 - p) with p the survival probability
 * 1677721 / 2^24 of the (0,1] uniform. */
double tmc::expected_interactions(const tmc_params* p, uint64_t kcap)
{
    const double pr = 1677721.0 / 16777216.0;
    uint32_t k1 = 0, k2 = 0;
    double c1, c2;
    if (weight_cadence(p, 1u << 20, &k1, &k2)) {
        c1 = (double)k1;
        c2 = (double)k2;
    } else {
        /* albedo so close to 1 that the cadence is long: continuous estimate */
        const double a = (double)derived_albedo(p);
        c1 = std::log(0.001) / std::log(a);
        c2 = std::log(0.1) / std::log(a);
    }
    double mean = c1 + c2 * pr / (1.0 - pr);
    if (kcap && mean > (double)kcap)
        mean = (double)kcap;
    return mean < 1.0 ? 1.0 : mean;
}

/* Entries of the binary32 walk's weight table (0: the walk multiplies).
 * The table is built per workgroup when the launch starts, so small runs
 * multiply; it must span the first cycle and eight more (a photon survives
 * eight roulettes with probability 1e-8), or the lanes past its end would
 * run the arithmetic beside it on most steps.  Automatic for both samplers
 * (measured at default physics, 2^30 photons: analytic 157.0 -> 150.6 ms,
 * rejection 260.0 -> 257.1 ms, whose retry iterations read an entry they do
 * not use). */
uint32_t tmc::weight_table_entries(const tmc_params* p, uint32_t precision, uint64_t photons, uint32_t request)
{
    if (precision != TMC_PRECISION_FP32 || request == TMC_WEIGHT_TABLE_OFF)
        return 0;
    if (request)
        return std::min<uint32_t>(request, TMC_WALK_TABLE_MAX);
    if (photons < (1ull << 20))
        return 0;
    uint32_t k1 = 0, k2 = 0;
    if (!weight_cadence(p, TMC_WALK_TABLE_MAX, &k1, &k2) || (uint64_t)k1 + 8ull * k2 > TMC_WALK_TABLE_MAX)
        return 0;
    return (uint32_t)std::min<uint64_t>(TMC_WALK_TABLE_MAX, (uint64_t)k1 + 24ull * k2);
}

int tmc::plan_passes(const std::vector<PlanSet>& sets, uint64_t chunk_work, uint32_t max_passes,
                     const ShapeFn& shape_of, std::vector<PlanPass>* passes, std::vector<PlanChunk>* chunks)
{
    passes->clear();
    chunks->clear();
    const size_t nsets = sets.size();
    double work = 0.0;
    for (const PlanSet& s : sets)
        work += s.L * (double)s.photons;

    /* LDS size classes: the distinct shell counts, ascending, grouped by the
     * launch shape they allow */
    std::vector<uint32_t> svals;
    for (const PlanSet& s : sets)
        svals.push_back(s.shells);
    std::sort(svals.begin(), svals.end());
    svals.erase(std::unique(svals.begin(), svals.end()), svals.end());
    std::vector<double> swork(svals.size(), 0.0);
    for (const PlanSet& s : sets) {
        const size_t k = (size_t)(std::lower_bound(svals.begin(), svals.end(), s.shells) - svals.begin());
        swork[k] += s.L * (double)s.photons;
    }
    std::vector<PlanPass> groups;
    for (size_t k = 0; k < svals.size(); ++k) {
        PlanShape shape;
        const int rc = shape_of(svals[k], &shape);
        if (rc != TMC_OK)
            return rc;
        if (!groups.empty() && groups.back().shape.grid == shape.grid && groups.back().shape.block == shape.block) {
            groups.back().hi = svals[k];
            groups.back().work += swork[k];
        } else {
            PlanPass g;
            g.lo = g.hi = svals[k];
            g.shape = shape;
            g.work = swork[k];
            groups.push_back(g);
        }
    }
    /* a class with under a tenth of the work joins its neighbour with more
     * shells (whose slots hold both), and there are at most max_passes */
    while (groups.size() > 1) {
        size_t m = 0;
        for (size_t k = 1; k < groups.size(); ++k)
            if (groups[k].work < groups[m].work)
                m = k;
        if (groups[m].work >= 0.1 * work && groups.size() <= max_passes)
            break;
        const size_t lo_i = m + 1 < groups.size() ? m : m - 1, hi_i = lo_i + 1;
        PlanPass g = groups[hi_i];
        g.lo = groups[lo_i].lo;
        g.work += groups[lo_i].work;
        groups[lo_i] = g;
        groups.erase(groups.begin() + (ptrdiff_t)hi_i);
    }

    /* chunk queues, one per pass */
    struct Item {
        PlanChunk ch;
        double dur;
    };
    std::vector<std::vector<Item>> items(groups.size());
    for (PlanPass& g : groups) {
        const double waves = (double)g.shape.grid * (g.shape.block / 64u);
        g.chunk_work = chunk_work ? chunk_work : (uint64_t)std::max(65536.0, g.work / (16.0 * std::max(waves, 1.0)));
    }
    for (size_t i = 0; i < nsets; ++i) {
        const uint64_t N = sets[i].photons;
        if (!N)
            continue;
        size_t g = 0;
        while (sets[i].shells > groups[g].hi)
            ++g;
        const double L = sets[i].L;
        double per_lane = std::ceil((double)groups[g].chunk_work / (64.0 * L));
        per_lane = std::min(std::max(per_lane, 1.0), (double)(1u << 24));
        uint64_t C = 64u * (uint64_t)per_lane;
        /* keep the queue below 2^24 chunks */
        const uint64_t min_c = (N >> 24) + 1;
        if (C < min_c)
            C = (min_c + 63) / 64 * 64;
        if (C > (1ull << 31))
            C = 1ull << 31;
        for (uint64_t first = 0; first < N; first += C) {
            const uint64_t cnt = std::min(C, N - first);
            items[g].push_back({ { (uint32_t)i, (uint32_t)cnt, first }, (double)((cnt + 63) / 64) * L });
        }
    }
    for (auto& v : items)
        std::stable_sort(v.begin(), v.end(), [](const Item& x, const Item& y) { return x.dur > y.dur; });
    /* the pass with the longest chunk first, so that the launch sequence ends
     * on short chunks */
    std::vector<size_t> order(groups.size());
    for (size_t g = 0; g < order.size(); ++g)
        order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        const double dx = items[x].empty() ? 0.0 : items[x][0].dur, dy = items[y].empty() ? 0.0 : items[y][0].dur;
        return dx > dy;
    });
    for (size_t g : order) {
        if (items[g].size() > UINT32_MAX)
            return TMC_E_RANGE; /* a pass's queue length is a 32-bit kernel argument */
        PlanPass q = groups[g];
        q.first_chunk = chunks->size();
        q.nchunks = items[g].size();
        for (const Item& it : items[g])
            chunks->push_back(it.ch);
        passes->push_back(q);
    }
    return TMC_OK;
}

void tmc::Arena::init(uint64_t words)
{
    free_.clear();
    if (words)
        free_[0] = words;
}

bool tmc::Arena::alloc(uint64_t n, uint64_t* off)
{
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= n) {
            *off = it->first;
            const uint64_t rest = it->second - n, at = it->first + n;
            free_.erase(it);
            if (rest)
                free_[at] = rest;
            return true;
        }
    }
    return false;
}

void tmc::Arena::release(uint64_t off, uint64_t n)
{
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

uint64_t tmc::chunk_photons(double work, double L, uint64_t max_per_lane, uint64_t left)
{
    double per_lane = std::floor(work / (64.0 * L));
    per_lane = std::min(std::max(per_lane, 1.0), (double)max_per_lane);
    return std::max<uint64_t>(1, std::min<uint64_t>(left, 64u * (uint64_t)per_lane));
}

double tmc::backlog_share(uint64_t photons, double L, double cap)
{
    return std::min((double)photons * L, cap);
}

double tmc::serve_chunk_scale(double backlog, double base_work, uint32_t waves, double chunks_per_wave,
                              uint32_t max_scale)
{
    return std::min(std::max(1.0, backlog / (chunks_per_wave * base_work * (double)waves)), (double)max_scale);
}
