// pa_fastx.cpp -- FASTA / FASTQ parsing into residues + offsets.  See include/pa_fastx.h.
//
// The text is cut into chunks at record starts (FASTA: a '>' at the start of a line; FASTQ: an
// '@' line whose second next line starts with '+', which a quality line starting with '@'
// cannot satisfy, since the line two below it is a sequence line).  Every chunk is parsed on its
// own thread into local buffers, which are then concatenated in record order (also in parallel).
#include "pa_fastx.h"
#include "pa_plan.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

struct pa_fastx {
    int32_t format = 0;
    pa::FillVector<uint8_t> residues, qualities;
    pa::FillVector<uint64_t> offsets, name_offsets;
    pa::FillVector<char> names;
};

namespace {

struct Fault {
    uint64_t at = 0;
    std::string what;
};

// A chunk's records: residues, qualities and names are written into buffers of the chunk's
// byte size (an upper bound for each), left uninitialised until written.
struct Chunk {
    uint64_t begin = 0, end = 0;   // record starts
    pa::FillVector<uint8_t> residues, qualities;
    pa::FillVector<char> names;
    size_t n_res = 0, n_qual = 0, n_names = 0;
    std::vector<uint64_t> lengths, name_lengths;
    bool failed = false;
    Fault fault;
};

inline uint64_t line_end(const char *t, uint64_t pos, uint64_t end)
{
    const void *nl = std::memchr(t + pos, '\n', end - pos);
    return nl ? uint64_t(static_cast<const char *>(nl) - t) : end;
}

inline uint64_t trim_cr(const char *t, uint64_t lo, uint64_t hi)
{
    return (hi > lo && t[hi - 1] == '\r') ? hi - 1 : hi;
}

inline bool blank(const char *t, uint64_t lo, uint64_t hi)
{
    for (uint64_t i = lo; i < hi; ++i)
        if (t[i] != ' ' && t[i] != '\t' && t[i] != '\r') return false;
    return true;
}

inline uint8_t fold(uint8_t c, bool upper) { return (upper && c >= 'a' && c <= 'z') ? uint8_t(c - 32) : c; }

// Writes the residues of [lo, hi) without blanks (FASTA sequence lines) to d; returns how many.
// Lines without blanks, the usual case, are copied whole.
size_t take_residues(const char *t, uint64_t lo, uint64_t hi, bool upper, uint8_t *d)
{
    const char *p = t + lo;
    size_t n = size_t(hi - lo);
    if (n && p[n - 1] == '\r') --n;
    if (!std::memchr(p, ' ', n) && !std::memchr(p, '\t', n) && !std::memchr(p, '\r', n)) {
        std::memcpy(d, p, n);
        if (upper)
            for (size_t i = 0; i < n; ++i) d[i] = fold(d[i], true);
        return n;
    }
    uint8_t *const d0 = d;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t c = uint8_t(p[i]);
        *d = fold(c, upper);
        d += (c != ' ' && c != '\t' && c != '\r');
    }
    return size_t(d - d0);
}

void add_name(Chunk &c, const char *t, uint64_t lo, uint64_t hi)
{
    std::memcpy(c.names.data() + c.n_names, t + lo, hi - lo);
    c.n_names += hi - lo;
    c.name_lengths.push_back(hi - lo);
}

void parse_fasta(const char *t, Chunk &c, bool upper)
{
    uint64_t pos = c.begin;
    while (pos < c.end) {
        uint64_t le = line_end(t, pos, c.end);
        if (t[pos] == ';' || blank(t, pos, le)) { pos = le + 1; continue; }
        if (t[pos] != '>') {
            c.failed = true;
            c.fault = {pos, "sequence data before the first '>' header"};
            return;
        }
        add_name(c, t, pos + 1, trim_cr(t, pos + 1, le));
        const size_t before = c.n_res;
        pos = le + 1;
        while (pos < c.end && t[pos] != '>') {
            le = line_end(t, pos, c.end);
            if (t[pos] != ';') c.n_res += take_residues(t, pos, le, upper, c.residues.data() + c.n_res);
            pos = le + 1;
        }
        c.lengths.push_back(c.n_res - before);
    }
}

void parse_fastq(const char *t, Chunk &c, bool upper)
{
    uint64_t pos = c.begin;
    while (pos < c.end) {
        uint64_t le = line_end(t, pos, c.end);
        if (blank(t, pos, le)) { pos = le + 1; continue; }
        if (t[pos] != '@') {
            c.failed = true;
            c.fault = {pos, "expected an '@' record header"};
            return;
        }
        const uint64_t ne = trim_cr(t, pos + 1, le);
        const uint64_t sl = le + 1;   // sequence line
        if (sl >= c.end) {
            c.failed = true;
            c.fault = {pos, "record without a sequence line"};
            return;
        }
        const uint64_t se = line_end(t, sl, c.end);
        const uint64_t pl = se + 1;   // '+' line
        if (pl >= c.end || t[pl] != '+') {
            c.failed = true;
            c.fault = {pl < c.end ? pl : pos, "expected a '+' line after the sequence (multi-line FASTQ is not supported)"};
            return;
        }
        const uint64_t ql = std::min(line_end(t, pl, c.end) + 1, c.end);   // quality line (empty at the end)
        const uint64_t qe = line_end(t, ql, c.end);
        const uint64_t s_hi = trim_cr(t, sl, se), q_hi = trim_cr(t, ql, qe);
        if (q_hi - ql != s_hi - sl) {
            c.failed = true;
            c.fault = {ql, "quality line length differs from the sequence length"};
            return;
        }
        add_name(c, t, pos + 1, ne);
        uint8_t *d = c.residues.data() + c.n_res;
        for (uint64_t i = sl; i < s_hi; ++i) d[i - sl] = fold(uint8_t(t[i]), upper);
        c.n_res += s_hi - sl;
        std::memcpy(c.qualities.data() + c.n_qual, t + ql, q_hi - ql);
        c.n_qual += q_hi - ql;
        c.lengths.push_back(s_hi - sl);
        pos = qe + 1;
    }
}

// First record start at or after `pos` (a line start), or `end`.
uint64_t next_record(const char *t, uint64_t pos, uint64_t end, int format)
{
    while (pos < end) {
        if (format == PA_FASTX_FASTA) {
            if (t[pos] == '>') return pos;
        } else if (t[pos] == '@') {
            const uint64_t l1 = line_end(t, pos, end) + 1;
            const uint64_t l2 = l1 < end ? line_end(t, l1, end) + 1 : end;
            if (l2 < end && t[l2] == '+') return pos;
        }
        pos = line_end(t, pos, end) + 1;
    }
    return end;
}

unsigned parse_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

pa_status report(char *err, size_t err_len, pa_status st, const std::string &msg)
{
    if (err && err_len) std::snprintf(err, err_len, "%s", msg.c_str());
    return st;
}

} // namespace

extern "C" {

pa_status pa_fastx_parse(const char *text, uint64_t bytes, uint32_t flags, pa_fastx **out, char *err, size_t err_len)
{
    if (!out || (bytes && !text)) return report(err, err_len, PA_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (err && err_len) err[0] = 0;
    uint64_t first = 0;
    while (first < bytes && (text[first] == ' ' || text[first] == '\t' || text[first] == '\r' || text[first] == '\n'))
        ++first;
    auto rec = new (std::nothrow) pa_fastx();
    if (!rec) return report(err, err_len, PA_ERR_OUT_OF_MEMORY, "out of host memory");
    if (first == bytes) {   // no records
        rec->format = PA_FASTX_FASTA;
        rec->offsets.assign(1, 0);
        rec->name_offsets.assign(1, 0);
        *out = rec;
        return PA_OK;
    }
    if (text[first] != '>' && text[first] != '@') {
        delete rec;
        return report(err, err_len, PA_ERR_INVALID_ARGUMENT, "neither FASTA ('>') nor FASTQ ('@') at byte " +
                      std::to_string(first));
    }
    const int format = text[first] == '>' ? PA_FASTX_FASTA : PA_FASTX_FASTQ;
    rec->format = format;
    const bool upper = (flags & PA_FASTX_UPPERCASE) != 0;

    // chunk boundaries: nominal cuts moved forward to the next record start
    const uint64_t span = bytes - first;
    const unsigned nt = parse_threads();
    const uint64_t nchunks = std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(nt) * 4, span >> 20));
    std::vector<uint64_t> cut(nchunks + 1);
    cut[0] = first;
    cut[nchunks] = bytes;
    auto for_chunks = [&](auto body) {
        std::vector<std::thread> th;
        const unsigned n = unsigned(std::min<uint64_t>(nt, nchunks));
        for (unsigned k = 0; k < n; ++k)
            th.emplace_back([&, k] { for (uint64_t c = k; c < nchunks; c += n) body(c); });
        for (auto &x : th) x.join();
    };
    for_chunks([&](uint64_t c) {
        if (c == 0) return;
        uint64_t p = first + span * c / nchunks;
        while (p < bytes && text[p - 1] != '\n') ++p;   // next line start
        cut[c] = next_record(text, p, bytes, format);
    });
    for (uint64_t c = 1; c <= nchunks; ++c) cut[c] = std::max(cut[c], cut[c - 1]);
    std::vector<Chunk> chunks(nchunks);
    try {
        for (uint64_t c = 0; c < nchunks; ++c) {
            const size_t span_c = size_t(cut[c + 1] - cut[c]);
            chunks[c].begin = cut[c];
            chunks[c].end = cut[c + 1];
            chunks[c].residues.resize(span_c);
            chunks[c].names.resize(span_c);
            if (format == PA_FASTX_FASTQ) chunks[c].qualities.resize(span_c);
        }
    } catch (const std::bad_alloc &) {
        delete rec;
        return report(err, err_len, PA_ERR_OUT_OF_MEMORY, "out of host memory");
    }
    for_chunks([&](uint64_t c) {
        if (format == PA_FASTX_FASTA) parse_fasta(text, chunks[c], upper);
        else parse_fastq(text, chunks[c], upper);
    });
    for (const Chunk &c : chunks)
        if (c.failed) {
            delete rec;
            return report(err, err_len, PA_ERR_INVALID_ARGUMENT, c.fault.what + " at byte " + std::to_string(c.fault.at));
        }

    // concatenate in record order
    std::vector<uint64_t> res_at(nchunks + 1, 0), rec_at(nchunks + 1, 0), name_at(nchunks + 1, 0);
    for (uint64_t c = 0; c < nchunks; ++c) {
        res_at[c + 1] = res_at[c] + chunks[c].n_res;
        rec_at[c + 1] = rec_at[c] + chunks[c].lengths.size();
        name_at[c + 1] = name_at[c] + chunks[c].n_names;
    }
    const uint64_t n = rec_at[nchunks];
    try {
        rec->residues.resize(res_at[nchunks]);
        if (format == PA_FASTX_FASTQ) rec->qualities.resize(res_at[nchunks]);
        rec->offsets.resize(n + 1);
        rec->name_offsets.resize(n + 1);
        rec->names.resize(name_at[nchunks]);
    } catch (const std::bad_alloc &) {
        delete rec;
        return report(err, err_len, PA_ERR_OUT_OF_MEMORY, "out of host memory");
    }
    for_chunks([&](uint64_t c) {
        const Chunk &k = chunks[c];
        if (k.n_res) std::memcpy(rec->residues.data() + res_at[c], k.residues.data(), k.n_res);
        if (k.n_qual) std::memcpy(rec->qualities.data() + res_at[c], k.qualities.data(), k.n_qual);
        if (k.n_names) std::memcpy(rec->names.data() + name_at[c], k.names.data(), k.n_names);
        uint64_t o = res_at[c], no = name_at[c];
        for (size_t r = 0; r < k.lengths.size(); ++r) {
            o += k.lengths[r];
            no += k.name_lengths[r];
            rec->offsets[rec_at[c] + r + 1] = o;
            rec->name_offsets[rec_at[c] + r + 1] = no;
        }
    });
    rec->offsets[0] = 0;
    rec->name_offsets[0] = 0;
    *out = rec;
    return PA_OK;
}

pa_status pa_fastx_read(const char *path, uint32_t flags, pa_fastx **out, char *err, size_t err_len)
{
    if (!path || !out) return report(err, err_len, PA_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    // mapped read-only and parsed in place; read into memory where mapping is not possible
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return report(err, err_len, PA_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    struct stat sb;
    if (::fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode) && sb.st_size > 0) {
        void *m = ::mmap(nullptr, size_t(sb.st_size), PROT_READ, MAP_PRIVATE, fd, 0);
        if (m != MAP_FAILED) {
            ::close(fd);
            const pa_status st = pa_fastx_parse(static_cast<const char *>(m), uint64_t(sb.st_size), flags, out, err,
                                                err_len);
            ::munmap(m, size_t(sb.st_size));
            return st;
        }
    }
    std::vector<char> text;
    char buf[1 << 16];
    ssize_t got;
    while ((got = ::read(fd, buf, sizeof buf)) > 0) text.insert(text.end(), buf, buf + got);
    ::close(fd);
    if (got < 0) return report(err, err_len, PA_ERR_INVALID_ARGUMENT, std::string("cannot read ") + path);
    return pa_fastx_parse(text.data(), text.size(), flags, out, err, err_len);
}

pa_status pa_fastx_get(const pa_fastx *r, pa_fastx_view *v)
{
    if (!r || !v) return PA_ERR_INVALID_ARGUMENT;
    v->format = r->format;
    v->n_records = r->offsets.size() - 1;
    v->residues = r->residues.data();
    v->offsets = r->offsets.data();
    v->names = r->names.data();
    v->name_offsets = r->name_offsets.data();
    v->qualities = r->format == PA_FASTX_FASTQ ? r->qualities.data() : nullptr;
    return PA_OK;
}

pa_status pa_fastx_destroy(pa_fastx *r)
{
    delete r;
    return PA_OK;
}

} // extern "C"
