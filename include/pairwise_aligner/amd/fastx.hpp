// pairwise_aligner/amd/fastx.hpp -- FASTA / FASTQ records for compute() (header-only wrapper of
// pa_fastx.h).  The reference takes character ranges and reads no files; this is the step in
// front of compute() for sequences kept in files:
//
//     auto reads = seqan::pairwise_aligner::fastx_records::read("reads.fq");
//     auto refs = seqan::pairwise_aligner::fastx_records::read("refs.fa", true);   // fold a-z
//     auto results = aligner.compute(reads.sequences(), refs.sequences());
//
// Records are views into one parsed buffer, which lives as long as any copy of the object.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include <pa_fastx.h>

namespace seqan::pairwise_aligner {
inline namespace v1 {

class fastx_records {
public:
    // Reads a FASTA or FASTQ file (the format is taken from its first non-blank byte).
    static fastx_records read(std::string const &path, bool uppercase = false)
    {
        pa_fastx *h = nullptr;
        char err[256] = {};
        if (pa_fastx_read(path.c_str(), uppercase ? PA_FASTX_UPPERCASE : 0u, &h, err, sizeof err) != PA_OK)
            throw std::runtime_error(std::string{"pa_fastx_read: "} + err);
        return fastx_records{h};
    }

    // Parses FASTA or FASTQ text.
    static fastx_records parse(std::string_view text, bool uppercase = false)
    {
        pa_fastx *h = nullptr;
        char err[256] = {};
        if (pa_fastx_parse(text.data(), text.size(), uppercase ? PA_FASTX_UPPERCASE : 0u, &h, err, sizeof err) != PA_OK)
            throw std::runtime_error(std::string{"pa_fastx_parse: "} + err);
        return fastx_records{h};
    }

    bool is_fastq() const noexcept { return view_.format == PA_FASTX_FASTQ; }
    std::size_t size() const noexcept { return std::size_t(view_.n_records); }

    // residues of record i
    std::string_view operator[](std::size_t i) const noexcept
    {
        return {reinterpret_cast<char const *>(view_.residues) + view_.offsets[i],
                std::size_t(view_.offsets[i + 1] - view_.offsets[i])};
    }
    std::string_view name(std::size_t i) const noexcept
    {
        return {view_.names + view_.name_offsets[i], std::size_t(view_.name_offsets[i + 1] - view_.name_offsets[i])};
    }
    // FASTQ qualities of record i (empty for FASTA)
    std::string_view quality(std::size_t i) const noexcept
    {
        if (!view_.qualities) return {};
        return {reinterpret_cast<char const *>(view_.qualities) + view_.offsets[i],
                std::size_t(view_.offsets[i + 1] - view_.offsets[i])};
    }
    // every record's residues, the form compute() takes for a bulk
    std::vector<std::string_view> sequences() const
    {
        std::vector<std::string_view> out(size());
        for (std::size_t i = 0; i < out.size(); ++i) out[i] = (*this)[i];
        return out;
    }

private:
    explicit fastx_records(pa_fastx *h) : handle_{h, [](pa_fastx *p) { pa_fastx_destroy(p); }}
    {
        pa_fastx_get(h, &view_);
    }

    std::shared_ptr<pa_fastx> handle_;
    pa_fastx_view view_{};
};

} // namespace v1
} // namespace seqan::pairwise_aligner
