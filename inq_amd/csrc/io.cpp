// io.cpp -- the reference's restart format for states, fields and arrays
// (src/operations/io.hpp:41-211, src/systems/electrons.hpp:323-381):
// raw native-endian files without headers, one per state column or field
// component.
//
//   states      dir/<ist>_<ispinor>   N complex128 real-space values, point
//               order (i0, i1, i2)     (io::save of an orbital_set, io.hpp:168-183)
//   field set   dir/<icomp>            N doubles per component (io.hpp:124-140)
//   array       dir/array              n doubles (io.hpp:104-120)
//
// <k> is the index printed with "%010ld" (utils::num_to_str, src/utils/num_str.hpp:20-24).
// Sphere states are written through the sphere -> real-space transform
// (to_real, x 1/N) and read back through the masked forward transform, so a
// state on the G sphere survives a save/load round trip up to FFT rounding.
#include "internal.hpp"

#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

namespace inqx {
namespace {

std::string num_str(long num) {
  char buf[12];
  std::snprintf(buf, 11, "%010ld", num);
  return std::string(buf);
}

void make_dir(const std::string &dir) {
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  INQX_CHECK(!ec, INQX_ERR_ARG, "cannot create directory '" + dir + "': " + ec.message());
}

void write_raw(const std::string &path, const void *data, size_t bytes) {
  std::FILE *f = std::fopen(path.c_str(), "wb");
  INQX_CHECK(f, INQX_ERR_ARG, "cannot create restart file '" + path + "'");
  const size_t w = std::fwrite(data, 1, bytes, f);
  const int c = std::fclose(f);
  INQX_CHECK(w == bytes && c == 0, INQX_ERR_ARG, "cannot write restart file '" + path + "'");
}

void read_raw(const std::string &path, void *data, size_t bytes) {
  std::FILE *f = std::fopen(path.c_str(), "rb");
  INQX_CHECK(f, INQX_ERR_ARG, "cannot open restart file '" + path + "'");
  const size_t r = std::fread(data, 1, bytes, f);
  std::fclose(f);
  INQX_CHECK(r == bytes, INQX_ERR_ARG, "restart file '" + path + "' is shorter than " + std::to_string(bytes) + " bytes");
}

std::string state_file(const std::string &dir, long ist) { return dir + "/" + num_str(ist) + "_" + num_str(0); }

// states per transform chunk: at most 256 MiB of complex128 real-space values
int state_chunk(long N, int nst) { return (int)std::max<long>(1, std::min<long>(nst, (256L << 20) / (16 * N))); }

void check_dir(const char *dir) { INQX_CHECK(dir && dir[0], INQX_ERR_ARG, "empty directory name"); }

}  // namespace
}  // namespace inqx

using namespace inqx;

extern "C" {

int inqx_states_save(inqx_ham h, inqx_basis bs, int nst, int precision, const void *psi, const char *dirname,
                     int first_state) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && bs && psi && nst > 0 && first_state >= 0, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(precision == INQX_FP64 || precision == INQX_FP32, INQX_ERR_ARG, "bad precision");
  check_dir(dirname);
  const std::string dir(dirname);
  make_dir(dir);
  const long N = h->g.N, npw = bs->b.npw;
  const size_t cs = precision == INQX_FP64 ? 16 : 8;
  const int chunk = state_chunk(N, nst);
  DevMem rs;
  rs.ensure(cs * (size_t)chunk * N);
  std::vector<char> host(cs * (size_t)chunk * N);
  std::vector<double> c128(precision == INQX_FP32 ? 2 * (size_t)N : 0);
  for (int c0 = 0; c0 < nst; c0 += chunk) {
    const int nc = std::min(chunk, nst - c0);
    const int st = inqx_sphere_to_real(h, bs, nc, precision, (const char *)psi + cs * (size_t)c0 * npw,
                                       INQX_STATE_MAJOR, rs.p);
    INQX_CHECK(st == INQX_OK, st, ctx->c.err);
    INQX_HIP(hipMemcpyAsync(host.data(), rs.p, cs * (size_t)nc * N, hipMemcpyDeviceToHost, ctx->c.stream));
    INQX_HIP(hipStreamSynchronize(ctx->c.stream));
    for (int i = 0; i < nc; ++i) {
      const char *src = host.data() + cs * (size_t)i * N;
      if (precision == INQX_FP64) {
        write_raw(state_file(dir, (long)first_state + c0 + i), src, 16 * (size_t)N);
      } else {
        const float *f = reinterpret_cast<const float *>(src);
        for (size_t k = 0; k < 2 * (size_t)N; ++k) c128[k] = (double)f[k];
        write_raw(state_file(dir, (long)first_state + c0 + i), c128.data(), 16 * (size_t)N);
      }
    }
  }
  API_END(ctx)
}

int inqx_states_load(inqx_ham h, inqx_basis bs, int nst, int precision, const char *dirname, int first_state,
                     void *psi) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && bs && psi && nst > 0 && first_state >= 0, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(precision == INQX_FP64 || precision == INQX_FP32, INQX_ERR_ARG, "bad precision");
  check_dir(dirname);
  const std::string dir(dirname);
  const long N = h->g.N, npw = bs->b.npw;
  const size_t cs = precision == INQX_FP64 ? 16 : 8;
  const int chunk = state_chunk(N, nst);
  DevMem rs;
  rs.ensure(cs * (size_t)chunk * N);
  std::vector<char> host(cs * (size_t)chunk * N);
  std::vector<double> c128(precision == INQX_FP32 ? 2 * (size_t)N : 0);
  for (int c0 = 0; c0 < nst; c0 += chunk) {
    const int nc = std::min(chunk, nst - c0);
    for (int i = 0; i < nc; ++i) {
      char *dst = host.data() + cs * (size_t)i * N;
      if (precision == INQX_FP64) {
        read_raw(state_file(dir, (long)first_state + c0 + i), dst, 16 * (size_t)N);
      } else {
        read_raw(state_file(dir, (long)first_state + c0 + i), c128.data(), 16 * (size_t)N);
        float *f = reinterpret_cast<float *>(dst);
        for (size_t k = 0; k < 2 * (size_t)N; ++k) f[k] = (float)c128[k];
      }
    }
    INQX_HIP(hipMemcpyAsync(rs.p, host.data(), cs * (size_t)nc * N, hipMemcpyHostToDevice, ctx->c.stream));
    const int st = inqx_real_to_sphere(h, bs, nc, precision, rs.p, INQX_STATE_MAJOR, 0,
                                       (char *)psi + cs * (size_t)c0 * npw);
    INQX_CHECK(st == INQX_OK, st, ctx->c.err);
    INQX_HIP(hipStreamSynchronize(ctx->c.stream));  // the host buffer is reused by the next chunk
  }
  API_END(ctx)
}

int inqx_field_save(inqx_ctx ctx, const char *dirname, long npoints, int ncomp, const double *v) {
  API_BEGIN
  INQX_CHECK(npoints > 0 && ncomp > 0 && v, INQX_ERR_ARG, "bad arguments");
  check_dir(dirname);
  const std::string dir(dirname);
  make_dir(dir);
  std::vector<double> col(npoints);
  for (int c = 0; c < ncomp; ++c) {
    for (long p = 0; p < npoints; ++p) col[p] = v[p * ncomp + c];
    write_raw(dir + "/" + num_str(c), col.data(), sizeof(double) * (size_t)npoints);
  }
  API_END(ctx)
}

int inqx_field_load(inqx_ctx ctx, const char *dirname, long npoints, int ncomp, double *v) {
  API_BEGIN
  INQX_CHECK(npoints > 0 && ncomp > 0 && v, INQX_ERR_ARG, "bad arguments");
  check_dir(dirname);
  const std::string dir(dirname);
  std::vector<double> col(npoints);
  for (int c = 0; c < ncomp; ++c) {
    read_raw(dir + "/" + num_str(c), col.data(), sizeof(double) * (size_t)npoints);
    for (long p = 0; p < npoints; ++p) v[p * ncomp + c] = col[p];
  }
  API_END(ctx)
}

int inqx_array_save(inqx_ctx ctx, const char *dirname, long n, const double *v) {
  API_BEGIN
  INQX_CHECK(n > 0 && v, INQX_ERR_ARG, "bad arguments");
  check_dir(dirname);
  make_dir(dirname);
  write_raw(std::string(dirname) + "/array", v, sizeof(double) * (size_t)n);
  API_END(ctx)
}

int inqx_array_load(inqx_ctx ctx, const char *dirname, long n, double *v) {
  API_BEGIN
  INQX_CHECK(n > 0 && v, INQX_ERR_ARG, "bad arguments");
  check_dir(dirname);
  read_raw(std::string(dirname) + "/array", v, sizeof(double) * (size_t)n);
  API_END(ctx)
}

}  // extern "C"
