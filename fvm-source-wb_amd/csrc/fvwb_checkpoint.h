// fvwb_checkpoint.h — the binary checkpoint format shared by the handles.
//
// A checkpoint is the device state that a time loop carries from one step to
// the next, written byte for byte, so a handle that loads it continues exactly
// as the one that saved it would have.  Layout (little endian, as the host):
//   char     magic[8]    "FVWBCKPT"
//   uint32   version     2 (2: the 2D FV state holds the four variables' rows
//                        of a grid row side by side; version 1 files, with
//                        separate planes, are rejected)
//   uint32   kind        which handle type wrote it (Kind below)
//   uint64   pbytes      size of the parameter block that follows
//   uint32   nseg        number of state segments
//   int32    parity      the handle's step parity
//   pbytes   the handle's parameter struct, as given to its create call
//   nseg times: uint64 bytes, then the segment's bytes
// A load checks magic, version, kind, the parameter block (byte for byte),
// every segment size against the handle it is loaded into and the file's
// length, all before its first copy to the device: a foreign, truncated or
// padded file is rejected with the device state untouched.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fvwb.h"

namespace fvwb_ckpt {

enum Kind : uint32_t { FV2D = 1, FV1D = 2, DG2D = 3, DG1D = 4, RD = 5 };

struct Segment {
  void* dev;     // device memory
  size_t bytes;
};

constexpr char kMagic[8] = {'F', 'V', 'W', 'B', 'C', 'K', 'P', 'T'};
constexpr uint32_t kVersion = 2;
constexpr size_t kChunk = (size_t)64 << 20;   // host staging per copy

inline int save(const char* path, Kind kind, const void* params, uint64_t pbytes, int32_t parity,
                const Segment* seg, uint32_t nseg, hipStream_t s, std::string& err) {
  if (!path) { err = "NULL path"; return FVWB_ERR_ARG; }
  if (hipStreamSynchronize(s) != hipSuccess) { err = "hipStreamSynchronize failed"; return FVWB_ERR_HIP; }
  FILE* f = std::fopen(path, "wb");
  if (!f) { err = std::string("cannot open ") + path + " for writing"; return FVWB_ERR_ARG; }
  const uint32_t version = kVersion, k = kind;
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&version, 4, 1, f) == 1 &&
            std::fwrite(&k, 4, 1, f) == 1 && std::fwrite(&pbytes, 8, 1, f) == 1 &&
            std::fwrite(&nseg, 4, 1, f) == 1 && std::fwrite(&parity, 4, 1, f) == 1 &&
            std::fwrite(params, 1, pbytes, f) == pbytes;
  std::vector<char> buf;
  for (uint32_t i = 0; ok && i < nseg; ++i) {
    const uint64_t n = seg[i].bytes;
    ok = std::fwrite(&n, 8, 1, f) == 1;
    for (size_t o = 0; ok && o < n; o += kChunk) {
      const size_t c = n - o < kChunk ? n - o : kChunk;
      buf.resize(c);
      if (hipMemcpy(buf.data(), (const char*)seg[i].dev + o, c, hipMemcpyDeviceToHost) != hipSuccess) {
        std::fclose(f);
        err = "hipMemcpy from the device failed";
        return FVWB_ERR_HIP;
      }
      ok = std::fwrite(buf.data(), 1, c, f) == c;
    }
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) { err = std::string("writing ") + path + " failed"; return FVWB_ERR_ARG; }
  return FVWB_OK;
}

inline int load(const char* path, Kind kind, const void* params, uint64_t pbytes, int32_t* parity,
                const Segment* seg, uint32_t nseg, hipStream_t s, std::string& err) {
  if (!path) { err = "NULL path"; return FVWB_ERR_ARG; }
  FILE* f = std::fopen(path, "rb");
  if (!f) { err = std::string("cannot open ") + path; return FVWB_ERR_ARG; }
  auto bad = [&](const std::string& why) {
    std::fclose(f);
    err = std::string(path) + ": " + why;
    return FVWB_ERR_ARG;
  };
  char magic[8];
  uint32_t version = 0, k = 0, ns = 0;
  uint64_t pb = 0;
  int32_t par = 0;
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0) return bad("not a checkpoint");
  if (std::fread(&version, 4, 1, f) != 1 || version != kVersion) return bad("unknown checkpoint version");
  if (std::fread(&k, 4, 1, f) != 1 || k != (uint32_t)kind) return bad("written by another handle type");
  if (std::fread(&pb, 8, 1, f) != 1 || pb != pbytes) return bad("parameter block of another size");
  if (std::fread(&ns, 4, 1, f) != 1 || ns != nseg) return bad("segment count differs");
  if (std::fread(&par, 4, 1, f) != 1 || (par != 0 && par != 1)) return bad("bad parity");
  std::vector<char> pblock(pb);
  if (std::fread(pblock.data(), 1, pb, f) != pb) return bad("truncated");
  if (std::memcmp(pblock.data(), params, pb) != 0) return bad("written by a handle with other parameters");
  // walk the segment headers and the file's end before touching the device
  const long data_start = std::ftell(f);
  if (data_start < 0) return bad("cannot position in the file");
  for (uint32_t i = 0; i < nseg; ++i) {
    uint64_t n = 0;
    if (std::fread(&n, 8, 1, f) != 1) return bad("truncated");
    if (n != seg[i].bytes) return bad("segment size differs");
    if (std::fseek(f, (long)n, SEEK_CUR) != 0) return bad("truncated");
  }
  const long data_end = std::ftell(f);
  if (data_end < 0 || std::fseek(f, 0, SEEK_END) != 0) return bad("cannot position in the file");
  const long file_end = std::ftell(f);
  if (file_end < data_end) return bad("truncated");
  if (file_end > data_end) return bad("trailing bytes after the last segment");
  if (std::fseek(f, data_start, SEEK_SET) != 0) return bad("cannot position in the file");
  std::vector<char> buf;
  for (uint32_t i = 0; i < nseg; ++i) {
    uint64_t n = 0;
    if (std::fread(&n, 8, 1, f) != 1 || n != seg[i].bytes) return bad("segment size differs");
    for (size_t o = 0; o < n; o += kChunk) {
      const size_t c = n - o < kChunk ? n - o : kChunk;
      buf.resize(c);
      if (std::fread(buf.data(), 1, c, f) != c) return bad("truncated");
      if (hipMemcpyAsync((char*)seg[i].dev + o, buf.data(), c, hipMemcpyHostToDevice, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        std::fclose(f);
        err = "hipMemcpy to the device failed";
        return FVWB_ERR_HIP;
      }
    }
  }
  std::fclose(f);
  *parity = par;
  return FVWB_OK;
}

}  // namespace fvwb_ckpt
