// CSV -> f64 ingest on the device: construct_file_iterator_skip_newline::<f64>
// (database/src/client.rs:163-182) over a file's bytes resident in HBM.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace buffgpu {

constexpr uint32_t kCsvBlock = 4096;  // bytes per workgroup (256 lanes x 16)

// The 128-bit truncation of 5^q, normalized to [2^127, 2^128): 5^q ~ (hi:lo) * 2^e.
struct Pow5 {
  uint64_t hi, lo;
  int32_t e;
  uint32_t exact;  // (hi:lo) * 2^e == 5^q exactly
};
constexpr int kPow5Min = -342, kPow5Max = 308;

// Per-block statistics (k_csv_blocks) and the prefix values of the blocks
// before it (k_csv_scan).
struct CsvBlock {
  uint32_t ntok;      // items ending in the block (delimiters, and the end of the text)
  uint32_t since;     // items ending after the block's last newline (all when it has none)
  uint32_t has_nl;
  uint32_t nonascii;
  int64_t last_delim; // absolute position of the block's last delimiter, -1 if none
};
struct CsvPre {
  uint64_t tok_base;  // items ending in earlier blocks
  uint64_t ok_base;   // parsed items of earlier blocks (after the parse)
  uint32_t line_idx;  // index within its line of the first item ending in the block
  uint32_t ok;        // parsed items of this block (written by the parse)
  int64_t prev_delim; // last delimiter before the block, -1 if none
};
// An item whose rounding the 19-digit path leaves open: its bytes and its slot.
struct CsvSlow {
  uint64_t start, end, slot;
};
// Totals written by the kernels: [0] items, [1] non-ASCII bytes seen, [2]
// parsed items, [3] items queued for the exact path, [4] items that path
// cannot take (too many digits, queue full), [5] the first such item's offset.
constexpr uint32_t kCsvTotals = 6;

uint64_t csv_scratch_bytes(uint64_t nblocks);
hipError_t launch_csv_count(const uint8_t* text, uint64_t len, uint8_t delim, uint8_t* scratch, uint64_t nblocks,
                            hipStream_t s);
hipError_t launch_csv_parse(const uint8_t* text, uint64_t len, uint8_t delim, uint32_t skip, const Pow5* tab,
                            uint8_t* scratch, uint64_t nblocks, double* dst, uint8_t* okb, CsvSlow* slow,
                            uint64_t slow_cap, hipStream_t s);
hipError_t launch_csv_slow(const uint8_t* text, const CsvSlow* slow, uint64_t count, uint8_t* scratch,
                           uint64_t nblocks, double* dst, hipStream_t s);
hipError_t launch_csv_compact(const uint8_t* scratch, uint64_t nblocks, const uint8_t* okb, const double* src,
                              double* out, hipStream_t s);
// totals pointer inside the scratch
uint64_t csv_totals_offset(uint64_t nblocks);
void csv_pow5_table(Pow5* out);  // host: the 651 entries, q = -342 .. 308

}  // namespace buffgpu
