/* pa_fastx.h -- FASTA / FASTQ records into the packed layout of the alignment entry points.
 *
 * The reference takes sequences as character ranges and reads no files (its only data file is
 * the FASTQ fixture test/data/in.fastq, declared in test/data/datasources.cmake:6-8); its
 * benchmarks generate sequences in memory (alignment_benchmark_fixture.hpp:130-139).  This is
 * the step in front of pa_score_batch / pa_batch_create for callers whose pairs come from files:
 * records are parsed on up to 16 host threads into one residue buffer plus n + 1 offsets, the
 * form those entry points take (seq = residues, off = offsets).
 *
 * Formats (detected from the first non-blank byte):
 *   FASTA  '>' header line, then sequence lines up to the next '>' line.  Blank lines and lines
 *          starting with ';' are skipped; spaces, tabs and carriage returns inside sequence
 *          lines are dropped.
 *   FASTQ  four-line records: '@' name, sequence, '+' (optionally the name again), qualities of
 *          the sequence's length.  Carriage returns are dropped; blank lines between records
 *          are skipped.  Multi-line FASTQ records are rejected.
 * A record's name is its header line after the marker, up to the end of the line.  Residue
 * bytes are kept as they are unless PA_FASTX_UPPERCASE folds a-z to A-Z (soft-masked input).
 * Host code only: no device is needed. */
#ifndef PA_FASTX_H
#define PA_FASTX_H

#include <stddef.h>
#include <stdint.h>

#include "pa_device.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { PA_FASTX_FASTA = 1, PA_FASTX_FASTQ = 2 };
enum { PA_FASTX_UPPERCASE = 1u };

typedef struct pa_fastx pa_fastx;

typedef struct pa_fastx_view {
    int32_t format;               /* PA_FASTX_FASTA or PA_FASTX_FASTQ */
    uint64_t n_records;
    const uint8_t *residues;      /* record r: residues[offsets[r], offsets[r + 1]) */
    const uint64_t *offsets;      /* n_records + 1 entries, offsets[0] = 0 */
    const char *names;            /* record r: names[name_offsets[r], name_offsets[r + 1]) */
    const uint64_t *name_offsets; /* n_records + 1 entries */
    const uint8_t *qualities;     /* FASTQ: the same offsets as residues; NULL for FASTA */
} pa_fastx_view;

/* Parses `bytes` bytes of FASTA or FASTQ text.  On failure *out is NULL and, if err is not
 * NULL, a message naming the byte offset of the fault is written into err (err_len bytes). */
pa_status pa_fastx_parse(const char *text, uint64_t bytes, uint32_t flags, pa_fastx **out, char *err,
                         size_t err_len);
/* The same for a file (mapped read-only, or read into memory where it cannot be mapped). */
pa_status pa_fastx_read(const char *path, uint32_t flags, pa_fastx **out, char *err, size_t err_len);
/* The records; the pointers stay valid until pa_fastx_destroy. */
pa_status pa_fastx_get(const pa_fastx *records, pa_fastx_view *view);
pa_status pa_fastx_destroy(pa_fastx *records);

#ifdef __cplusplus
}
#endif

#endif
