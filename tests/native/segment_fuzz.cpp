// Host-only harness: the segment parser (thindb_amd/csrc/segment.cpp) under
// AddressSanitizer and UBSan, fed randomly corrupted and truncated copies of
// one valid segment file.  The library's error hook and the HBM staging call
// are stubbed: only the parsing, decompression and int_streams restore run.
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <random>
#include "../../include/thindb_gpu.h"
namespace tdbg { namespace rt {
int32_t fail(int32_t code, const char* fmt, ...) { return code; }
}}
extern "C" int32_t tdbg_column_stage_encoded(tdbg_context*, const tdbg_column_key*, uint32_t, uint32_t, const void*, uint64_t, uint64_t, uint32_t, const tdbg_block_stats*) { return 0; }
int main(int argc, char** argv) {
  FILE* f = fopen(argv[1], "rb"); std::vector<uint8_t> img; int c; while ((c = fgetc(f)) != EOF) img.push_back((uint8_t)c); fclose(f);
  std::mt19937_64 rng(1);
  int ok = 0, bad = 0;
  const int trials = argc > 2 ? atoi(argv[2]) : 20000;
  for (int t = 0; t < trials; t++) {
    std::vector<uint8_t> m = img;
    if (t % 5 == 4) m.resize(rng() % m.size());
    else for (int k = 0, n = 1 + rng() % 8; k < n; k++) m[rng() % m.size()] = (uint8_t)rng();
    std::vector<uint8_t> exact(m.begin(), m.end());  // exact-size heap copy so ASan sees overruns
    uint8_t* buf = (uint8_t*)malloc(exact.size() ? exact.size() : 1); memcpy(buf, exact.data(), exact.size());
    tdbg_segment* s = nullptr;
    if (tdbg_segment_open(buf, exact.size(), 2, &s) == 0) {
      tdbg_segment_info inf; tdbg_segment_get_info(s, &inf);
      for (uint32_t rg = 0; rg < inf.n_row_groups && rg < 64; rg++)
        for (uint32_t col = 0; col < 2; col++) {
          tdbg_segment_block b;
          if (tdbg_segment_block_info(s, rg, col, &b)) continue;
          if (b.payload_bytes > (64u << 20)) continue;
          std::vector<uint8_t> out(b.payload_bytes ? b.payload_bytes : 1);
          int rc = tdbg_segment_read_block(s, rg, col, out.data(), b.payload_bytes);
          rc ? bad++ : ok++;
          tdbg_segment_stage((tdbg_context*)1, s, 1, rg, col, 5);
        }
      tdbg_segment_close(s);
    } else bad++;
    free(buf);
  }
  printf("ok %d bad %d\n", ok, bad);
}
