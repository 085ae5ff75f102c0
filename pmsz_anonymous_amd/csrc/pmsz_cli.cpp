// pmsz_cli.cpp -- the command-line drop-in for the reference executable.
//
//   pmsz "<path>,<W>,<H>,<D>" <abs_eb> <sz3|zfp|dq> <PX> <PY> <PZ> [options]
//   pmsz <path> <W> <H> <D> <abs_eb> <sz3|zfp|dq> <PX> <PY> <PZ> [options]
//
// The positional contract is the one main() parses (pMSz.cu:2625-2667): a raw
// little-endian field (x fastest), the absolute error bound, the compressor id
// and the subdomain grid.  The reference reads float64 only (pMSz.cu:2821);
// --dtype f32 reads float32.  Output lines keep the reference's keys
// ("iteration is:", pMSz.cu:2073; "total storage overhead is:", "original CR
// is:", "overall CR is:", "overall edit ratio is: ... total_edited:",
// pMSz.cu:3140-3144).  Like the reference (gather_cost_sparse,
// pMSz.cu:2429-2543) it writes the edit side channel to <prefix>_deltas.bin and
// <prefix>_edits.bin, prefix "" by default (the working directory).
//
// options: --dtype f32|f64   --mode lazy|sync|naive   --max-sweeps N
//          --clamp reference|bound-safe   --ghost reference|full   --q N
//          --out <path> (corrected field, raw)   --symbols <path> (delta symbols)
//          --edits-prefix <prefix>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pmsz.h"
#include "../../include/pmsz_io.h"

namespace {

int usage() {
  std::fprintf(stderr,
               "usage: pmsz \"<path>,<W>,<H>,<D>\" <abs_eb> <sz3|zfp|dq> <PX> <PY> <PZ> "
               "[--dtype f32|f64] [--mode lazy|sync|naive] [--max-sweeps N] "
               "[--clamp reference|bound-safe] [--ghost reference|full] [--q N] "
               "[--out path] [--symbols path] [--edits-prefix prefix]\n");
  return 2;
}

bool parse_dims(const std::string &s, std::string &path, int64_t &w, int64_t &h, int64_t &d) {
  size_t c1 = s.find(',');
  if (c1 == std::string::npos) return false;
  path = s.substr(0, c1);
  std::istringstream iss(s.substr(c1 + 1));
  char a, b;
  if (!(iss >> w >> a >> h >> b >> d) || a != ',' || b != ',') return false;
  return w > 0 && h > 0 && d > 0;
}

}  // namespace

// The README's form, "<path> <W> <H> <D> <abs_eb> <compressor> <PX> <PY> <PZ>"
// (README.md:68-74): the first argument names a file without the ",W,H,D"
// suffix and the next three are whole numbers.
bool readme_form(int argc, char **argv) {
  if (argc < 10 || std::strchr(argv[1], ',')) return false;
  for (int k = 2; k <= 4; k++) {
    char *end = nullptr;
    const long long v = std::strtoll(argv[k], &end, 10);
    if (!end || *end || v <= 0) return false;
  }
  return true;
}

int main(int argc, char **argv) {
  if (argc < 7) return usage();
  std::string path;
  int64_t W = 0, H = 0, D = 0;
  // the code's form (pMSz.cu:2625-2667) or the README's, told apart as above
  const bool rf = readme_form(argc, argv);
  const int a = rf ? 3 : 0;  // shift of the arguments after the extent
  if (rf) {
    path = argv[1];
    W = std::atoll(argv[2]);
    H = std::atoll(argv[3]);
    D = std::atoll(argv[4]);
  } else if (!parse_dims(argv[1], path, W, H, D)) {
    std::fprintf(stderr, "Parsing error for dimensions\n");
    return 1;
  }
  pmsz_params p;
  pmsz_default_params(&p);
  p.dtype = PMSZ_F64;
  p.nx = W;
  p.ny = H;
  p.nz = D;
  p.abs_eb = std::strtod(argv[2 + a], nullptr);
  std::string comp = argv[3 + a];
  p.sub[0] = std::atoi(argv[4 + a]);
  p.sub[1] = std::atoi(argv[5 + a]);
  p.sub[2] = std::atoi(argv[6 + a]);
  std::string out_path, sym_path, edits_prefix;
  for (int i = 7 + a; i < argc; i++) {
    std::string o = argv[i];
    auto val = [&](void) -> std::string {
      if (i + 1 >= argc) {
        usage();
        std::exit(2);
      }
      return argv[++i];
    };
    if (o == "--dtype") {
      std::string v = val();
      p.dtype = v == "f32" ? PMSZ_F32 : PMSZ_F64;
    } else if (o == "--mode") {
      const std::string v = val();
      p.mode = v == "sync" ? PMSZ_SYNC : v == "naive" ? PMSZ_NAIVE : PMSZ_LAZY;
    } else if (o == "--max-sweeps") {
      p.max_sweeps = std::atoi(val().c_str());
    } else if (o == "--clamp") {
      p.clamp_mode = val() == "reference" ? PMSZ_CLAMP_REFERENCE : PMSZ_CLAMP_BOUND_SAFE;
    } else if (o == "--ghost") {
      p.ghost_policy = val() == "reference" ? PMSZ_GHOST_REFERENCE : PMSZ_GHOST_FULL;
    } else if (o == "--q") {
      p.q = std::atoi(val().c_str());
    } else if (o == "--out") {
      out_path = val();
    } else if (o == "--symbols") {
      sym_path = val();
    } else if (o == "--edits-prefix") {
      edits_prefix = val();
    } else {
      return usage();
    }
  }
  if (comp != "sz3" && comp != "zfp" && comp != "dq") {
    std::fprintf(stderr, "unknown compressor %s (sz3|zfp|dq)\n", comp.c_str());
    return 1;
  }
  const int64_t n = W * H * D;
  const size_t es = p.dtype == PMSZ_F32 ? 4 : 8;
  std::vector<unsigned char> f(n * es), g0(n * es), g(n * es);
  if (pmsz_io_read_raw(path.c_str(), p.dtype, n, f.data()) != PMSZ_IO_OK) {
    std::fprintf(stderr, "cannot read %lld values from %s\n", (long long)n, path.c_str());
    return 1;
  }
  void *df = nullptr, *dg0 = nullptr, *dg = nullptr;
  uint8_t *dsym = nullptr;
  if (hipMalloc(&df, f.size()) != hipSuccess || hipMalloc(&dg0, g.size()) != hipSuccess ||
      hipMalloc(&dg, g.size()) != hipSuccess || hipMalloc((void **)&dsym, n) != hipSuccess) {
    std::fprintf(stderr, "device allocation failed\n");
    return 1;
  }
  if (hipMemcpy(df, f.data(), f.size(), hipMemcpyHostToDevice) != hipSuccess) {
    std::fprintf(stderr, "host to device copy failed\n");
    return 1;
  }
  auto t0 = std::chrono::steady_clock::now();
  int rc;
  int64_t cbytes = 0;
  double est = 0;
  const char *size_kind = "stream";  // how cbytes was obtained
  if (comp == "dq") {
    // the GPU front end: dual-quantisation Lorenzo, core-only per rank block,
    // Huffman-coded into a stream on the device and decoded back to check it
    int64_t cap = 0;
    void *dstream = nullptr;
    void *dback = nullptr;
    pmsz_codec_info ci;
    if (pmsz_codec_dq_stream_bound(&p, &cap) != PMSZ_OK || hipMalloc(&dstream, cap) != hipSuccess ||
        hipMalloc(&dback, n * es) != hipSuccess) {
      std::fprintf(stderr, "codec buffers failed\n");
      return 1;
    }
    rc = pmsz_codec_dq_compress(&p, df, dg0, (uint8_t *)dstream, cap, &cbytes, &ci);
    if (rc == PMSZ_OK) rc = pmsz_codec_dq_decompress((const uint8_t *)dstream, cbytes, dback);
    if (rc != PMSZ_OK) {
      std::fprintf(stderr, "dq codec: %s (%d)\n", pmsz_strerror(rc), rc);
      return 1;
    }
    std::vector<unsigned char> back(n * es);
    if (hipMemcpy(g0.data(), dg0, g0.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(back.data(), dback, back.size(), hipMemcpyDeviceToHost) != hipSuccess)
      return 1;
    if (std::memcmp(back.data(), g0.data(), back.size()) != 0) {
      std::fprintf(stderr, "dq codec: the decoded stream differs from g0\n");
      return 1;
    }
    std::printf("dq stream: %lld bytes (entropy estimate %lld), decoded back exactly\n",
                (long long)cbytes, (long long)ci.compressed_bytes);
    (void)hipFree(dstream);
    (void)hipFree(dback);
  } else if (comp == "sz3") {
    int64_t un = 0;
    rc = pmsz_io_szlike(p.dtype, W, H, D, f.data(), p.abs_eb, g0.data(), &un, &est);
    cbytes = (int64_t)est;
    size_kind = "entropy estimate";  // the host SZ-like stand-in writes no stream
  } else {
    rc = pmsz_io_zfp(p.dtype, W, H, D, f.data(), p.abs_eb, g0.data(), &cbytes);
  }
  if (rc != PMSZ_IO_OK) {
    std::fprintf(stderr, "compressor %s failed (%d)\n", comp.c_str(), rc);
    return 1;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::memcpy(g.data(), g0.data(), g.size());
  if (hipMemcpy(dg0, g0.data(), g0.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dg, g0.data(), g0.size(), hipMemcpyHostToDevice) != hipSuccess) {
    std::fprintf(stderr, "host to device copy failed\n");
    return 1;
  }
  std::printf("topology preserving started!\n");
  pmsz_stats st;
  auto t2 = std::chrono::steady_clock::now();
  rc = pmsz_correct(&p, df, dg, nullptr, &st);
  auto t3 = std::chrono::steady_clock::now();
  if (rc != PMSZ_OK && rc != PMSZ_NOT_CONVERGED) {
    std::fprintf(stderr, "pmsz_correct: %s (%d)", pmsz_strerror(rc), rc);
    if (rc == PMSZ_EDEVICE) std::fprintf(stderr, ": %s", hipGetErrorString(hipGetLastError()));
    std::fprintf(stderr, "\n");
    return 1;
  }
  std::printf("iteration is: %lld\n", (long long)st.ite);
  // the edit side channel: packed symbols and clamp values on the device, zstd
  // frames per rank block on the host (gather_cost_sparse, pMSz.cu:2447-2523)
  auto t4 = std::chrono::steady_clock::now();
  int64_t cap = 0;
  void *dpk = nullptr, *dvals = nullptr;
  if (pmsz_edit_packed_bytes(&p, &cap) != PMSZ_OK || hipMalloc(&dpk, cap + 16) != hipSuccess ||
      hipMalloc(&dvals, n * es) != hipSuccess) {
    std::fprintf(stderr, "edit buffers failed\n");
    return 1;
  }
  const int nblocks = p.sub[0] * p.sub[1] * p.sub[2];
  std::vector<pmsz_edit_block> blocks(nblocks);
  pmsz_edit_info info;
  rc = pmsz_encode_edits(&p, df, dg0, dg, (uint8_t *)dpk, dvals, blocks.data(), &info);
  if (rc != PMSZ_OK) {
    std::fprintf(stderr, "pmsz_encode_edits: %s (%d)\n", pmsz_strerror(rc), rc);
    return 1;
  }
  std::vector<uint8_t> packed(cap + 16);
  std::vector<unsigned char> vals((size_t)std::max<int64_t>(info.values, 1) * es);
  if (hipMemcpy(packed.data(), dpk, cap, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(vals.data(), dvals, info.values * es, hipMemcpyDeviceToHost) != hipSuccess)
    return 1;
  auto t5 = std::chrono::steady_clock::now();
  std::printf("[Time] extract_edits: %.6f s\n", std::chrono::duration<double>(t5 - t4).count());
  const int64_t edited = info.edited;
  int64_t nd = 0, ne = 0;
  int wrc = pmsz_io_write_edits(edits_prefix.c_str(), p.dtype, nblocks, blocks.data(),
                                packed.data(), vals.data(), 5, 3, &nd, &ne);
  auto t6 = std::chrono::steady_clock::now();
  if (wrc != PMSZ_IO_OK) {
    std::fprintf(stderr, "cannot write the edit files (%d)\n", wrc);
    return 1;
  }
  std::printf("[Time] compress: %.6f s\n", std::chrono::duration<double>(t6 - t5).count());
  std::printf("compressed deltas: %lld, %lld\n", (long long)nd, (long long)ne);
  if (!sym_path.empty() && pmsz_edit_symbols(&p, df, dg0, dg, dsym, nullptr) != PMSZ_OK) {
    std::fprintf(stderr, "edit extraction failed\n");
    return 1;
  }
  if (hipMemcpy(g.data(), dg, g.size(), hipMemcpyDeviceToHost) != hipSuccess) return 1;
  if (!out_path.empty() && pmsz_io_write_raw(out_path.c_str(), p.dtype, n, g.data()) != 0) {
    std::fprintf(stderr, "cannot write %s\n", out_path.c_str());
    return 1;
  }
  if (!sym_path.empty()) {
    std::vector<uint8_t> sym(n);
    if (hipMemcpy(sym.data(), dsym, n, hipMemcpyDeviceToHost) != hipSuccess) return 1;
    FILE *fp = std::fopen(sym_path.c_str(), "wb");
    if (!fp || std::fwrite(sym.data(), 1, n, fp) != (size_t)n) return 1;
    std::fclose(fp);
  }
  double max_err = 0;
  int64_t viol = 0;
  pmsz_io_max_abs_err(p.dtype, n, f.data(), g.data(), p.abs_eb, &max_err, &viol);
  double secs = std::chrono::duration<double>(t3 - t2).count();
  std::printf("compression time: %.6f s (%s, %lld bytes)\n",
              std::chrono::duration<double>(t1 - t0).count(), comp.c_str(), (long long)cbytes);
  std::printf("comfputational overhead is: %.6f\n", secs);
  std::printf("sweeps: %lld edit, %lld check, outer iterations %lld, status %s\n",
              (long long)st.edit_sweeps, (long long)st.check_sweeps, (long long)st.outer_iters,
              rc == PMSZ_OK ? "converged" : "NOT CONVERGED");
  std::printf("max abs error: %.17g (bound %.17g, violations %lld)\n", max_err, p.abs_eb,
              (long long)viol);
  // pMSz.cu:3133-3144: ratios over the input file's size
  const double orig = (double)n * (double)es, overhead = (double)(nd + ne);
  std::printf("total storage overhead is: %lld\n", (long long)(nd + ne));
  std::printf("compressed size: %lld bytes (%s)\n", (long long)cbytes, size_kind);
  std::printf("original CR is: %.17g\n", cbytes > 0 ? orig / (double)cbytes : 0.0);
  std::printf("overall CR is: %.17g\n", orig / ((double)cbytes + overhead));
  std::printf("overall edit ratio is: %.17g total_edited:%lld\n", (double)edited / (double)n,
              (long long)edited);
  std::printf("throughput: %.3f GVoxels/s (device %.3f ms)\n", n / secs / 1e9,
              st.t_total_s * 1e3);
  (void)hipFree(df);
  (void)hipFree(dg0);
  (void)hipFree(dg);
  (void)hipFree(dsym);
  (void)hipFree(dpk);
  (void)hipFree(dvals);
  return rc == PMSZ_OK ? 0 : 3;
}
