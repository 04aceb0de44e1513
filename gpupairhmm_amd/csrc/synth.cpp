// Seeded synthetic workloads.
//
// Every record is generated from its own RNG stream (seeded from the
// workload seed and the record index), so a workload is identical however
// many threads build it.  Common model (all configurations):
//   - a record's reference is uniform random ACGT;
//   - each haplotype is the reference mutated with 1% SNPs and 0.2% indels of
//     1-5 bp (insertion or deletion, equally likely), then trimmed or
//     extended with random bases to its target length;
//   - each read is drawn from a random haplotype of the record at a random
//     offset (from the reference when no haplotype is long enough), with
//     per-base substitution errors;
//   - base quality is uniform phred in [qlo, qhi]; insertion and deletion
//     qualities are 'I' (40) except a trailing 'N' (45); gcp is '+' (10).
// Configurations (counts are records x reads x haps):
//   0: 1,000 x 8 x 4; reads 100 bp; haps 120 bp; Q 10-40; 1% errors
//   1: 10,000 x 25 x 4; reads 150 bp; haps U[150,350]; Q 10-40; 1% errors
//   2: 20,000 x U[1,500] x U[1,64]; read length lognormal(median 150,
//      sigma 0.35) clipped to [36,300]; hap length = the record's median read
//      length + lognormal(median 100, sigma 0.5), clipped to <= 1000
//   3: 50,000 x 16 x 8; reads 250 bp; haps U[2000,5000]; Q 2-20; 5% errors
//   4: serving requests: the configuration-2 record distribution
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>

#include "host_batch.hpp"

namespace phmm {
namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        for (auto& v : s)
            v = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()  // xoshiro256**
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    int range(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
    double normal()
    {
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    double lognormal(double median, double sigma) { return median * std::exp(sigma * normal()); }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

std::string random_seq(Rng& rng, int n)
{
    std::string s((size_t)n, 'A');
    for (auto& c : s)
        c = kBases[rng.next() & 3];
    return s;
}

char other_base(Rng& rng, char b)
{
    for (;;) {
        const char c = kBases[rng.next() & 3];
        if (c != b)
            return c;
    }
}

std::string mutate(Rng& rng, const std::string& ref, double snp, double indel)
{
    std::string out;
    out.reserve(ref.size() + 16);
    for (size_t i = 0; i < ref.size(); ++i) {
        const double u = rng.uniform();
        if (u < indel) {
            const int n = rng.range(1, 5);
            if (rng.next() & 1) {
                out.push_back(ref[i]);
                out += random_seq(rng, n);
            } else {
                i += (size_t)n - 1;  // delete ref[i .. i+n-1]
            }
        } else if (u < indel + snp) {
            out.push_back(other_base(rng, ref[i]));
        } else {
            out.push_back(ref[i]);
        }
    }
    return out;
}

std::string fit_length(Rng& rng, std::string s, int n)
{
    if ((int)s.size() > n)
        s.resize((size_t)n);
    else if ((int)s.size() < n)
        s += random_seq(rng, n - (int)s.size());
    return s;
}

struct Spec {
    int records;
    int qlo, qhi;
    double err;
};

Spec spec_of(int config)
{
    switch (config) {
    case 0: return {1000, 10, 40, 0.01};
    case 1: return {10000, 10, 40, 0.01};
    case 2: return {20000, 10, 40, 0.01};
    case 3: return {50000, 2, 20, 0.05};
    default: return {20000, 10, 40, 0.01};
    }
}

SequenceGroup make_record(int config, const Spec& sp, uint64_t seed, int index)
{
    uint64_t x = seed * 0x100000001b3ULL + (uint64_t)index * 0x9e3779b97f4a7c15ULL + (uint64_t)config;
    Rng rng(Rng::splitmix(x));
    int R = 0, H = 0;
    std::vector<int> rl, hl;
    switch (config) {
    case 0:
        R = 8;
        H = 4;
        rl.assign(R, 100);
        hl.assign(H, 120);
        break;
    case 1:
        R = 25;
        H = 4;
        rl.assign(R, 150);
        for (int h = 0; h < H; ++h)
            hl.push_back(rng.range(150, 350));
        break;
    case 3:
        R = 16;
        H = 8;
        rl.assign(R, 250);
        for (int h = 0; h < H; ++h)
            hl.push_back(rng.range(2000, 5000));
        break;
    default: {  // 2 and 4
        R = rng.range(1, 500);
        H = rng.range(1, 64);
        for (int r = 0; r < R; ++r) {
            const int L = (int)std::lround(rng.lognormal(150.0, 0.35));
            rl.push_back(std::min(300, std::max(36, L)));
        }
        std::vector<int> sorted = rl;
        std::nth_element(sorted.begin(), sorted.begin() + R / 2, sorted.end());
        const int med = sorted[(size_t)R / 2];
        for (int h = 0; h < H; ++h) {
            const int extra = (int)std::lround(rng.lognormal(100.0, 0.5));
            hl.push_back(std::max(1, std::min(1000, med + extra)));
        }
        break;
    }
    }
    const int maxh = *std::max_element(hl.begin(), hl.end());
    const int maxr = *std::max_element(rl.begin(), rl.end());
    const std::string ref = random_seq(rng, std::max(maxh, maxr) + 16);
    std::vector<std::string> haps;
    for (int h = 0; h < H; ++h)
        haps.push_back(fit_length(rng, mutate(rng, ref.substr(0, (size_t)hl[h] + 8), 0.01, 0.002),
                                  hl[h]));

    SequenceGroup g;
    g.num_reads = R;
    g.num_haps = H;
    for (int r = 0; r < R; ++r) {
        const int L = rl[r];
        std::string src;
        const int hsel = rng.range(0, H - 1);
        if ((int)haps[hsel].size() >= L) {
            const int off = rng.range(0, (int)haps[hsel].size() - L);
            src = haps[hsel].substr((size_t)off, (size_t)L);
        } else {
            const int off = rng.range(0, (int)ref.size() - L);
            src = ref.substr((size_t)off, (size_t)L);
        }
        for (auto& c : src)
            if (rng.uniform() < sp.err)
                c = other_base(rng, c);
        const int pad = padding_of(L);
        for (int i = 0; i < L; ++i) {
            g.read_data.push_back((uint8_t)src[i]);
            g.base_quals.push_back((uint8_t)rng.range(sp.qlo, sp.qhi));
            g.ins_quals.push_back((uint8_t)(i + 1 < L ? 40 : 45));
            g.del_quals.push_back((uint8_t)(i + 1 < L ? 40 : 45));
            g.gcp_quals.push_back(10);
        }
        for (int i = 0; i < pad; ++i) {
            g.read_data.push_back('N');
            g.base_quals.push_back(45);
            g.ins_quals.push_back(45);
            g.del_quals.push_back(45);
            g.gcp_quals.push_back(45);
        }
        g.read_lengths.push_back(L);
    }
    for (int h = 0; h < H; ++h) {
        const std::string& s = haps[h];
        g.hap_data.insert(g.hap_data.end(), s.begin(), s.end());
        g.hap_data.insert(g.hap_data.end(), (size_t)padding_of((int)s.size()), 'N');
        g.hap_lengths.push_back((int)s.size());
    }
    return g;
}

}  // namespace

void synth_generate(int config, uint64_t seed, int32_t num_records, HostBatch& out)
{
    const Spec sp = spec_of(config);
    const int n = num_records >= 0 ? num_records : sp.records;
    std::vector<SequenceGroup> groups((size_t)n);
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 32u));
    if ((unsigned)n < 64)
        nt = 1;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            for (int i = (int)t; i < n; i += (int)nt)
                groups[(size_t)i] = make_record(config, sp, seed, i);
        });
    for (auto& th : pool)
        th.join();
    out.assign(groups);
}

}  // namespace phmm
