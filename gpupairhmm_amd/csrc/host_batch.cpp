// Host batch: parser, writer and batch assembly (see host_batch.hpp).
#include "host_batch.hpp"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "common.hpp"

namespace phmm {

void HostBatch::clear()
{
    read_data.clear();
    hap_data.clear();
    base_quals.clear();
    ins_quals.clear();
    del_quals.clear();
    gcp_quals.clear();
    read_lengths.clear();
    hap_lengths.clear();
    read_offsets.clear();
    hap_offsets.clear();
    num_reads.clear();
    num_haps.clear();
    reads_prefix.clear();
    haps_prefix.clear();
    num_alignments = 0;
}

phmm_batch_view HostBatch::view() const
{
    phmm_batch_view v{};
    v.num_groups = (int32_t)num_reads.size();
    v.num_reads = (int32_t)read_lengths.size();
    v.num_haps = (int32_t)hap_lengths.size();
    v.read_bytes = (int64_t)read_data.size();
    v.hap_bytes = (int64_t)hap_data.size();
    v.num_alignments = num_alignments;
    v.read_data = read_data.data();
    v.hap_data = hap_data.data();
    v.base_quals = base_quals.data();
    v.ins_quals = ins_quals.data();
    v.del_quals = del_quals.data();
    v.gcp_quals = gcp_quals.data();
    v.read_lengths = read_lengths.data();
    v.hap_lengths = hap_lengths.data();
    v.read_offsets = read_offsets.data();
    v.hap_offsets = hap_offsets.data();
    v.num_reads_per_group = num_reads.data();
    v.num_haps_per_group = num_haps.data();
    v.reads_prefix = reads_prefix.data();
    v.haps_prefix = haps_prefix.data();
    return v;
}

void HostBatch::assign(const std::vector<SequenceGroup>& groups)
{
    clear();
    size_t rb = 0, hb = 0, nr = 0, nh = 0;
    for (const auto& g : groups) {
        rb += g.read_data.size();
        hb += g.hap_data.size();
        nr += g.read_lengths.size();
        nh += g.hap_lengths.size();
    }
    if (rb > (size_t)INT32_MAX || hb > (size_t)INT32_MAX)
        throw std::runtime_error("batch exceeds 2^31 bytes of sequence data");
    read_data.reserve(rb);
    base_quals.reserve(rb);
    ins_quals.reserve(rb);
    del_quals.reserve(rb);
    gcp_quals.reserve(rb);
    hap_data.reserve(hb);
    read_lengths.reserve(nr);
    hap_lengths.reserve(nh);
    auto app = [](std::vector<uint8_t>& dst, const std::vector<uint8_t>& src) {
        dst.insert(dst.end(), src.begin(), src.end());
    };
    for (const auto& g : groups) {
        app(read_data, g.read_data);
        app(hap_data, g.hap_data);
        app(base_quals, g.base_quals);
        app(ins_quals, g.ins_quals);
        app(del_quals, g.del_quals);
        app(gcp_quals, g.gcp_quals);
        read_lengths.insert(read_lengths.end(), g.read_lengths.begin(), g.read_lengths.end());
        hap_lengths.insert(hap_lengths.end(), g.hap_lengths.begin(), g.hap_lengths.end());
        num_reads.push_back(g.num_reads);
        num_haps.push_back(g.num_haps);
    }
    read_offsets.resize(read_lengths.size() + 1);
    read_offsets[0] = 0;
    for (size_t i = 0; i < read_lengths.size(); ++i)
        read_offsets[i + 1] = read_offsets[i] + padded_length(read_lengths[i]);
    hap_offsets.resize(hap_lengths.size() + 1);
    hap_offsets[0] = 0;
    for (size_t i = 0; i < hap_lengths.size(); ++i)
        hap_offsets[i + 1] = hap_offsets[i] + padded_length(hap_lengths[i]);
    reads_prefix.resize(num_reads.size() + 1);
    haps_prefix.resize(num_haps.size() + 1);
    reads_prefix[0] = haps_prefix[0] = 0;
    num_alignments = 0;
    for (size_t g = 0; g < num_reads.size(); ++g) {
        reads_prefix[g + 1] = reads_prefix[g] + num_reads[g];
        haps_prefix[g + 1] = haps_prefix[g] + num_haps[g];
        num_alignments += (int64_t)num_reads[g] * num_haps[g];
    }
}

// stoi-like: skip leading whitespace, optional sign, digits; trailing text is
// ignored as std::stoi ignores it (file_parser_worker.cu:53-54).
static bool parse_int(const std::string& s, size_t from, long* out)
{
    const char* b = s.c_str() + from;
    char* e = nullptr;
    errno = 0;
    const long v = std::strtol(b, &e, 10);
    if (e == b || errno == ERANGE)
        return false;
    *out = v;
    return true;
}

SequenceGroup parse_group(std::istream& in, const std::string& header, int64_t* line_no)
{
    const size_t split = header.find(' ');
    long nr = 0, nh = 0;
    if (split == std::string::npos || !parse_int(header, 0, &nr) || !parse_int(header, split, &nh))
        throw std::runtime_error("line " + std::to_string(*line_no) +
                                 ": malformed record header '" + header + "'");
    if (nr < 0 || nh < 0 || nr > INT32_MAX / 2 || nh > INT32_MAX / 2)
        throw std::runtime_error("line " + std::to_string(*line_no) +
                                 ": invalid read/hap count in header");
    SequenceGroup g;
    g.num_reads = (int)nr;
    g.num_haps = (int)nh;
    constexpr uint8_t pad_char = 'N';
    constexpr int qoff = -33;
    std::string line;
    for (long i = 0; i < nr; ++i) {
        if (!std::getline(in, line))
            throw std::runtime_error("unexpected end of input inside a record (read " +
                                     std::to_string(i) + ")");
        ++*line_no;
        const size_t sp = line.find(' ');
        if (sp == std::string::npos)
            throw std::runtime_error("line " + std::to_string(*line_no) +
                                     ": read line without quality fields");
        const int L = (int)sp;
        if (line.size() < (size_t)5 * L + 4)
            throw std::runtime_error("line " + std::to_string(*line_no) +
                                     ": read line shorter than 5 fields of length " +
                                     std::to_string(L));
        const int pad = padding_of(L);
        const uint8_t* lb = reinterpret_cast<const uint8_t*>(line.data());
        g.read_data.insert(g.read_data.end(), lb, lb + L);
        g.read_data.insert(g.read_data.end(), pad, pad_char);
        std::vector<uint8_t>* qv[4] = {&g.base_quals, &g.ins_quals, &g.del_quals, &g.gcp_quals};
        for (int k = 0; k < 4; ++k) {
            const uint8_t* f = lb + (size_t)(k + 1) * (L + 1);
            for (int x = 0; x < L; ++x)
                qv[k]->push_back((uint8_t)(f[x] + qoff));
            for (int x = 0; x < pad; ++x)
                qv[k]->push_back((uint8_t)(pad_char + qoff));
        }
        g.read_lengths.push_back(L);
    }
    for (long i = 0; i < nh; ++i) {
        if (!std::getline(in, line))
            throw std::runtime_error("unexpected end of input inside a record (hap " +
                                     std::to_string(i) + ")");
        ++*line_no;
        const int H = (int)line.size();
        const uint8_t* lb = reinterpret_cast<const uint8_t*>(line.data());
        g.hap_data.insert(g.hap_data.end(), lb, lb + H);
        g.hap_data.insert(g.hap_data.end(), padding_of(H), pad_char);
        g.hap_lengths.push_back(H);
    }
    return g;
}

std::string validate_host_view(const phmm_batch_view& v)
{
    if (v.num_groups < 0 || v.num_reads < 0 || v.num_haps < 0)
        return "negative counts";
    if (v.num_groups > 0 && (!v.num_reads_per_group || !v.num_haps_per_group || !v.reads_prefix ||
                             !v.haps_prefix))
        return "missing group arrays";
    if (v.num_reads > 0 && (!v.read_lengths || !v.read_offsets || !v.read_data || !v.base_quals ||
                            !v.ins_quals || !v.del_quals || !v.gcp_quals))
        return "missing read arrays";
    if (v.num_haps > 0 && (!v.hap_lengths || !v.hap_offsets || !v.hap_data))
        return "missing hap arrays";
    if (v.num_groups > 0) {
        if (v.reads_prefix[0] != 0 || v.haps_prefix[0] != 0)
            return "reads_prefix[0] and haps_prefix[0] must be 0";
        int64_t na = 0;
        for (int g = 0; g < v.num_groups; ++g) {
            if (v.num_reads_per_group[g] < 0 || v.num_haps_per_group[g] < 0)
                return "negative group size";
            if (v.reads_prefix[g + 1] != v.reads_prefix[g] + v.num_reads_per_group[g] ||
                v.haps_prefix[g + 1] != v.haps_prefix[g] + v.num_haps_per_group[g])
                return "group prefix sums inconsistent at group " + std::to_string(g);
            na += (int64_t)v.num_reads_per_group[g] * v.num_haps_per_group[g];
        }
        if (v.reads_prefix[v.num_groups] != v.num_reads || v.haps_prefix[v.num_groups] != v.num_haps)
            return "group prefix sums do not match read/hap counts";
        if (na != v.num_alignments)
            return "num_alignments does not match the groups";
    } else if (v.num_alignments != 0) {
        return "alignments without groups";
    }
    if (v.num_alignments > INT32_MAX)
        return "more than 2^31-1 alignments in one batch";
    for (int i = 0; i < v.num_reads; ++i) {
        const int L = v.read_lengths[i];
        if (L < 0 || v.read_offsets[i] < 0 ||
            (int64_t)v.read_offsets[i] + padded_length(L) > v.read_bytes ||
            v.read_offsets[i + 1] < v.read_offsets[i] + padded_length(L) || (v.read_offsets[i] & 3))
            return "read " + std::to_string(i) + ": offsets/length outside the padded layout";
    }
    for (int i = 0; i < v.num_haps; ++i) {
        const int H = v.hap_lengths[i];
        if (H < 0 || v.hap_offsets[i] < 0 ||
            (int64_t)v.hap_offsets[i] + padded_length(H) > v.hap_bytes ||
            v.hap_offsets[i + 1] < v.hap_offsets[i] + padded_length(H) || (v.hap_offsets[i] & 3))
            return "hap " + std::to_string(i) + ": offsets/length outside the padded layout";
    }
    return std::string();
}

void format_scores(const phmm_batch_view& v, const float* scores, std::string& out)
{
    char buf[64];
    int64_t k = 0;
    for (int g = 0; g < v.num_groups; ++g) {
        const int R = v.num_reads_per_group[g], H = v.num_haps_per_group[g];
        int n = std::snprintf(buf, sizeof buf, "%d %d\n", R, H);
        out.append(buf, (size_t)n);
        for (int r = 0; r < R; ++r) {
            for (int h = 0; h < H; ++h) {
                if (h > 0)
                    out.push_back(' ');
                n = std::snprintf(buf, sizeof buf, "%.5f", (double)scores[k++]);
                out.append(buf, (size_t)n);
            }
            out.push_back('\n');
        }
    }
}

}  // namespace phmm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using phmm::HostBatch;

struct phmm_host_batch {
    HostBatch b;
};

extern "C" int phmm_host_batch_parse(const char* text, size_t len, phmm_host_batch** out)
{
    if (!out || (!text && len))
        return phmm::fail(PHMM_EINVAL, "phmm_host_batch_parse: null argument");
    *out = nullptr;
    try {
        std::istringstream in(std::string(text ? text : "", len));
        std::vector<phmm::SequenceGroup> groups;
        std::string line;
        int64_t line_no = 0;
        while (std::getline(in, line)) {
            ++line_no;
            if (line.empty())
                continue;
            groups.push_back(phmm::parse_group(in, line, &line_no));
        }
        auto* hb = new phmm_host_batch;
        hb->b.assign(groups);
        *out = hb;
        return PHMM_OK;
    } catch (const std::bad_alloc&) {
        return phmm::fail(PHMM_ENOMEM, "phmm_host_batch_parse: out of memory");
    } catch (const std::exception& e) {
        return phmm::fail(PHMM_EPARSE, e.what());
    }
}

extern "C" int phmm_parse_file(const char* path, int64_t file_batchsize, phmm_batch_callback cb,
                               void* user)
{
    if (!path || !cb)
        return phmm::fail(PHMM_EINVAL, "phmm_parse_file: null argument");
    std::ifstream in(path);
    if (!in)
        return phmm::fail(PHMM_EIO, std::string("Could not open file ") + path);
    if (file_batchsize <= 0)
        file_batchsize = INT64_MAX;
    try {
        std::vector<phmm::SequenceGroup> groups;
        HostBatch batch;
        int64_t n_align = 0, line_no = 0;
        std::string line;
        // A batch closes at the first record where the running count of
        // alignments reaches file_batchsize (file_parser_worker.cu:50-66).
        while (std::getline(in, line)) {
            ++line_no;
            if (line.empty())
                continue;
            groups.push_back(phmm::parse_group(in, line, &line_no));
            n_align += (int64_t)groups.back().num_reads * groups.back().num_haps;
            if (n_align >= file_batchsize) {
                batch.assign(groups);
                groups.clear();
                n_align = 0;
                const phmm_batch_view v = batch.view();
                const int rc = cb(&v, user);
                if (rc != 0)
                    return rc;
            }
        }
        if (!groups.empty()) {
            batch.assign(groups);
            const phmm_batch_view v = batch.view();
            const int rc = cb(&v, user);
            if (rc != 0)
                return rc;
        }
        return PHMM_OK;
    } catch (const std::bad_alloc&) {
        return phmm::fail(PHMM_ENOMEM, "phmm_parse_file: out of memory");
    } catch (const std::exception& e) {
        return phmm::fail(PHMM_EPARSE, e.what());
    }
}

extern "C" int phmm_host_batch_view(const phmm_host_batch* b, phmm_batch_view* out)
{
    if (!b || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_host_batch_view: null argument");
    *out = b->b.view();
    return PHMM_OK;
}

extern "C" void phmm_host_batch_free(phmm_host_batch* b) { delete b; }

extern "C" int phmm_synth_generate(int config, uint64_t seed, int32_t num_records,
                                   phmm_host_batch** out)
{
    if (!out)
        return phmm::fail(PHMM_EINVAL, "phmm_synth_generate: null argument");
    *out = nullptr;
    if (config < 0 || config > 4)
        return phmm::fail(PHMM_EINVAL, "phmm_synth_generate: config must be 0..4");
    try {
        auto* hb = new phmm_host_batch;
        phmm::synth_generate(config, seed, num_records, hb->b);
        *out = hb;
        return PHMM_OK;
    } catch (const std::bad_alloc&) {
        return phmm::fail(PHMM_ENOMEM, "phmm_synth_generate: out of memory");
    } catch (const std::exception& e) {
        return phmm::fail(PHMM_EINVAL, e.what());
    }
}

extern "C" int phmm_host_batch_to_text(const phmm_batch_view* v, char** text, size_t* len)
{
    if (!v || !text || !len)
        return phmm::fail(PHMM_EINVAL, "phmm_host_batch_to_text: null argument");
    const std::string err = phmm::validate_host_view(*v);
    if (!err.empty())
        return phmm::fail(PHMM_EINVAL, "phmm_host_batch_to_text: " + err);
    std::string s;
    for (int g = 0; g < v->num_groups; ++g) {
        const int R = v->num_reads_per_group[g], H = v->num_haps_per_group[g];
        s += std::to_string(R) + " " + std::to_string(H) + "\n";
        for (int r = 0; r < R; ++r) {
            const int i = v->reads_prefix[g] + r;
            const int L = v->read_lengths[i], o = v->read_offsets[i];
            s.append(reinterpret_cast<const char*>(v->read_data + o), (size_t)L);
            const uint8_t* qs[4] = {v->base_quals, v->ins_quals, v->del_quals, v->gcp_quals};
            for (int k = 0; k < 4; ++k) {
                s.push_back(' ');
                for (int x = 0; x < L; ++x)
                    s.push_back((char)(qs[k][o + x] + 33));
            }
            s.push_back('\n');
        }
        for (int h = 0; h < H; ++h) {
            const int i = v->haps_prefix[g] + h;
            s.append(reinterpret_cast<const char*>(v->hap_data + v->hap_offsets[i]),
                     (size_t)v->hap_lengths[i]);
            s.push_back('\n');
        }
    }
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    if (!p)
        return phmm::fail(PHMM_ENOMEM, "phmm_host_batch_to_text: out of memory");
    std::memcpy(p, s.data(), s.size());
    p[s.size()] = 0;
    *text = p;
    *len = s.size();
    return PHMM_OK;
}

extern "C" int phmm_format_scores(const phmm_batch_view* v, const float* scores, char** text,
                                  size_t* len)
{
    if (!v || !text || !len || (v->num_alignments > 0 && !scores))
        return phmm::fail(PHMM_EINVAL, "phmm_format_scores: null argument");
    std::string s;
    phmm::format_scores(*v, scores, s);
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    if (!p)
        return phmm::fail(PHMM_ENOMEM, "phmm_format_scores: out of memory");
    std::memcpy(p, s.data(), s.size());
    p[s.size()] = 0;
    *text = p;
    *len = s.size();
    return PHMM_OK;
}

extern "C" void phmm_free_text(char* text) { std::free(text); }

extern "C" int64_t phmm_batch_cells(const phmm_batch_view* v)
{
    if (!v)
        return -1;
    // sum over groups of (sum of read lengths) * (sum of hap lengths)
    // (computeNumberOfDPCells, gpuworker.cuh:495-520)
    int64_t total = 0;
    for (int g = 0; g < v->num_groups; ++g) {
        int64_t sr = 0, sh = 0;
        for (int r = 0; r < v->num_reads_per_group[g]; ++r)
            sr += v->read_lengths[v->reads_prefix[g] + r];
        for (int h = 0; h < v->num_haps_per_group[g]; ++h)
            sh += v->hap_lengths[v->haps_prefix[g] + h];
        total += sr * sh;
    }
    return total;
}

// One region (reads x haps with phred qualities) as one padded record of a
// host batch, scored by phmm_score_batch; log10 likelihoods widened to
// double for callers that keep them so.
extern "C" int phmm_compute_likelihoods(phmm_engine* eng, const phmm_region_read* reads,
                                        int32_t num_reads, const phmm_region_hap* haps,
                                        int32_t num_haps, double* out)
{
    if (!eng || num_reads < 0 || num_haps < 0 || (num_reads > 0 && !reads) ||
        (num_haps > 0 && !haps) || (num_reads > 0 && num_haps > 0 && !out))
        return phmm::fail(PHMM_EINVAL, "phmm_compute_likelihoods: bad arguments");
    if (num_reads == 0 || num_haps == 0)
        return PHMM_OK;
    constexpr uint8_t pad_char = 'N';
    constexpr uint8_t pad_qual = 'N' - 33;  // the parser's padding, as a phred value
    phmm::SequenceGroup g;
    g.num_reads = num_reads;
    g.num_haps = num_haps;
    for (int32_t i = 0; i < num_reads; ++i) {
        const phmm_region_read& r = reads[i];
        if (r.length < 0 || (r.length > 0 && (!r.bases || !r.base_quals || !r.ins_quals ||
                                              !r.del_quals || !r.gcp_quals)))
            return phmm::fail(PHMM_EINVAL, "phmm_compute_likelihoods: read " + std::to_string(i) +
                                               ": negative length or null array");
        const int pad = phmm::padding_of(r.length);
        g.read_data.insert(g.read_data.end(), r.bases, r.bases + r.length);
        g.read_data.insert(g.read_data.end(), pad, pad_char);
        const uint8_t* q[4] = {r.base_quals, r.ins_quals, r.del_quals, r.gcp_quals};
        std::vector<uint8_t>* qv[4] = {&g.base_quals, &g.ins_quals, &g.del_quals, &g.gcp_quals};
        for (int k = 0; k < 4; ++k) {
            qv[k]->insert(qv[k]->end(), q[k], q[k] + r.length);
            qv[k]->insert(qv[k]->end(), pad, pad_qual);
        }
        g.read_lengths.push_back(r.length);
    }
    for (int32_t i = 0; i < num_haps; ++i) {
        const phmm_region_hap& h = haps[i];
        if (h.length < 0 || (h.length > 0 && !h.bases))
            return phmm::fail(PHMM_EINVAL, "phmm_compute_likelihoods: hap " + std::to_string(i) +
                                               ": negative length or null bases");
        g.hap_data.insert(g.hap_data.end(), h.bases, h.bases + h.length);
        g.hap_data.insert(g.hap_data.end(), phmm::padding_of(h.length), pad_char);
        g.hap_lengths.push_back(h.length);
    }
    phmm::HostBatch hb;
    try {
        hb.assign({g});
    } catch (const std::exception& e) {
        return phmm::fail(PHMM_ENOMEM, std::string("phmm_compute_likelihoods: ") + e.what());
    }
    const phmm_batch_view v = hb.view();
    std::vector<float> scores((size_t)num_reads * (size_t)num_haps);
    const int rc = phmm_score_batch(eng, &v, scores.data());
    if (rc != PHMM_OK)
        return rc;
    for (size_t i = 0; i < scores.size(); ++i)
        out[i] = (double)scores[i];
    return PHMM_OK;
}
