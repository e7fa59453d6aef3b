// ddx core, part 7: the Pgx DDS-table .npy format (host only).
//
// A Pgx file is an int32 .npy array of shape (2, n, 4): keys[n][4], then values[n][4]
// (reference src/npy.h:1-63).
//   key[s] (Pgx suit s = S,H,D,C, i.e. our suit s): 2-bit owners, card k = A,2,3..K at
//          bit shift 2*(12-k); so rank r (0 = deuce .. 12 = ace) sits at k = (r + 1) % 13.
//   value[seat]: tricks of declarer `seat` in strain st (C,D,H,S,NT) at nibble 4*(4-st).
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

namespace ddx {

struct PgxFile {
  std::vector<int32_t> keys, values;  // n x 4 each
  long long n = 0;
};

// Reads rows [start, start + count) of a Pgx file; count < 0 = to the end. Returns false
// with a message in *err on any format problem.
inline bool pgx_read(const char* path, long long start, long long count, PgxFile* out,
                     std::string* err) {
  FILE* f = fopen(path, "rb");
  if (!f) {
    *err = std::string("cannot open ") + path;
    return false;
  }
  unsigned char pre[12];
  if (fread(pre, 1, 10, f) != 10 || memcmp(pre, "\x93NUMPY", 6) != 0) {
    fclose(f);
    *err = "not a .npy file";
    return false;
  }
  long hlen = pre[8] | (pre[9] << 8), off = 10;
  if (pre[6] >= 2) {
    if (fread(pre + 10, 1, 2, f) != 2) {
      fclose(f);
      *err = "truncated header";
      return false;
    }
    hlen |= ((long)pre[10] << 16) | ((long)pre[11] << 24);
    off = 12;
  }
  std::string hdr(hlen, '\0');
  if (fread(&hdr[0], 1, hlen, f) != (size_t)hlen) {
    fclose(f);
    *err = "truncated header";
    return false;
  }
  size_t sp = hdr.find("'shape': (2, ");
  if (sp == std::string::npos || hdr.find("<i4") == std::string::npos ||
      hdr.find("'fortran_order': False") == std::string::npos) {
    fclose(f);
    *err = "expected a C-order int32 array of shape (2, n, 4)";
    return false;
  }
  long long total = atoll(hdr.c_str() + sp + 13);
  if (start < 0) start += total;
  if (count < 0) count = total - start;
  if (start < 0 || start + count > total) {
    fclose(f);
    *err = "rows out of range";
    return false;
  }
  long data = off + hlen;
  out->n = count;
  out->keys.resize(count * 4);
  out->values.resize(count * 4);
  bool ok = fseek(f, data + start * 16, SEEK_SET) == 0 &&
            fread(out->keys.data(), 16, count, f) == (size_t)count &&
            fseek(f, data + (total + start) * 16, SEEK_SET) == 0 &&
            fread(out->values.data(), 16, count, f) == (size_t)count;
  fclose(f);
  if (!ok) *err = "truncated data";
  return ok;
}

// Hands of row i.
inline void pgx_hands(const PgxFile& d, long long i, uint64_t* h) {
  h[0] = h[1] = h[2] = h[3] = 0;
  for (int s = 0; s < 4; s++) {
    uint32_t key = (uint32_t)d.keys[i * 4 + s];
    for (int r = 0; r < 13; r++) {
      int k = (r + 1) % 13;
      h[(key >> (2 * (12 - k))) & 3] |= 1ull << (16 * s + r);
    }
  }
}

// Stored tricks of declarer `seat` in strain `st`.
inline int pgx_tricks(const PgxFile& d, long long i, int seat, int st) {
  return (int)(((uint32_t)d.values[i * 4 + seat] >> (4 * (4 - st))) & 15);
}

// Writes deals and their tables (20 per deal, [strain][declarer]) as a Pgx file, so that
// any Pgx reader (the reference's dd_cpu included) can load and check them.
inline bool pgx_write(const char* path, const uint64_t* hands, const uint8_t* tables, long long n) {
  FILE* f = fopen(path, "wb");
  if (!f) return false;
  char hdr[128];
  int len = snprintf(hdr, sizeof hdr, "{'descr': '<i4', 'fortran_order': False, 'shape': (2, %lld, 4), }", n);
  int total = 10 + len + 1;
  int pad = (64 - total % 64) % 64;
  std::string h(hdr, len);
  h.append(pad, ' ');
  h.push_back('\n');
  unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, (unsigned char)(h.size() & 255),
                           (unsigned char)(h.size() >> 8)};
  bool ok = fwrite(pre, 1, 10, f) == 10 && fwrite(h.data(), 1, h.size(), f) == h.size();
  std::vector<int32_t> keys(n * 4), vals(n * 4);
  for (long long i = 0; i < n; i++) {
    for (int s = 0; s < 4; s++) {
      uint32_t key = 0;
      for (int r = 0; r < 13; r++) {
        int seat = 0;
        for (int p = 0; p < 4; p++)
          if (hands[i * 4 + p] >> (16 * s + r) & 1) seat = p;
        key |= (uint32_t)seat << (2 * (12 - (r + 1) % 13));
      }
      keys[i * 4 + s] = (int32_t)key;
    }
    for (int seat = 0; seat < 4; seat++) {
      uint32_t v = 0;
      for (int st = 0; st < 5; st++) v |= (uint32_t)(tables ? tables[i * 20 + st * 4 + seat] : 0) << (4 * (4 - st));
      vals[i * 4 + seat] = (int32_t)v;
    }
  }
  ok = ok && fwrite(keys.data(), 4, keys.size(), f) == keys.size() &&
       fwrite(vals.data(), 4, vals.size(), f) == vals.size();
  fclose(f);
  return ok;
}

}  // namespace ddx
