// Interpolation grids on disk: a binary cache of a built grid (so that a restart loads instead of rebuilding) and the
// reference's cube export.
//
//   cube export   InterpolationEnergyGrid::writeOutput (src/raspakit/interpolation_energy_grid.cpp:4760-4790)
//   metadata      the grid's archive record (unit cell, origin, point counts, cell counts, order; :4793-4836); the
//                 reference does not store the values (src/raspakit/system.cpp:1382-1384 rebuilds them on restart),
//                 this cache stores them behind a key of the inputs they were built from.
#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rg_internal.hpp"

namespace
{

constexpr char kMagic[8] = {'R', 'G', 'G', 'R', 'I', 'D', '0', '1'};

struct CacheHeader
{
  char magic[8];
  uint64_t key;
  int32_t values, n[3], rect, pad;
  double cell[9];
  uint64_t count;
};
static_assert(sizeof(CacheHeader) == 8 + 8 + 24 + 72 + 8, "cache header layout");

// FNV-1a over raw bytes
struct Fnv
{
  uint64_t h = 1469598103934665603ull;
  void bytes(const void* p, size_t n)
  {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i)
    {
      h ^= c[i];
      h *= 1099511628211ull;
    }
  }
  template <typename T>
  void value(const T& v)
  {
    bytes(&v, sizeof(T));
  }
};

// Streams the grid through a host buffer of at most 64 MiB.
constexpr size_t kChunk = size_t(8) << 20;  // doubles

}  // namespace

extern "C"
{

rg_status rg_grid_key(const rg_forcefield* ff, const rg_framework* fw, rg_grid_type type, int probe_pseudo_atom,
                      double cutoff_supercell, int precision, double split_radius, unsigned long long* key)
{
  if (!ff || !fw || !key) return RG_ERR_INVALID;
  Fnv f;
  f.value(int(type));
  f.value(probe_pseudo_atom);
  f.value(cutoff_supercell);
  f.value(precision);
  f.value(precision == RG_GRID_SPLIT_FP32 ? split_radius : 0.0);
  f.bytes(fw->cell, sizeof(fw->cell));
  f.value(fw->rect);
  f.value(fw->n);
  f.bytes(fw->pos.data(), fw->pos.size() * sizeof(double));
  f.bytes(fw->type.data(), fw->type.size() * sizeof(int));
  f.bytes(fw->charge.data(), fw->charge.size() * sizeof(double));
  f.value(ff->n);
  f.bytes(ff->eps.data(), ff->eps.size() * sizeof(double));
  f.bytes(ff->sigma.data(), ff->sigma.size() * sizeof(double));
  f.bytes(ff->shift.data(), ff->shift.size() * sizeof(double));
  f.bytes(ff->lj.data(), ff->lj.size() * sizeof(int));
  f.value(ff->rc_framework);
  f.value(ff->rc_coulomb);
  f.value(ff->alpha);
  f.value(ff->overlap);
  *key = f.h;
  return RG_OK;
}

rg_status rg_grid_save(rg_ctx* ctx, const rg_grid* grid, unsigned long long key, const char* path)
{
  if (!ctx || !grid || !path) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_save: invalid arguments");
  RG_ENTER(ctx);
  FILE* f = std::fopen(path, "wb");
  if (!f) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_save: cannot open %s for writing", path);
  CacheHeader h{};
  std::memcpy(h.magic, kMagic, sizeof(kMagic));
  h.key = key;
  h.values = grid->values;
  for (int a = 0; a < 3; ++a) h.n[a] = grid->n[a];
  h.rect = grid->rect;
  std::memcpy(h.cell, grid->cell, sizeof(h.cell));
  h.count = grid->count;
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
  std::vector<double> buf(std::min(kChunk, grid->count));
  for (size_t off = 0; ok && off < grid->count; off += kChunk)
  {
    const size_t c = std::min(kChunk, grid->count - off);
    hipError_t e = hipMemcpyAsync(buf.data(), grid->d + off, c * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
    {
      std::fclose(f);
      return rg::hipfail(ctx, e, "rg_grid_save");
    }
    ok = std::fwrite(buf.data(), sizeof(double), c, f) == c;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_save: writing %s failed", path);
  return RG_OK;
}

rg_status rg_grid_load(rg_ctx* ctx, rg_grid* grid, unsigned long long key, const char* path)
{
  if (!ctx || !grid || !path) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_load: invalid arguments");
  RG_ENTER(ctx);
  FILE* f = std::fopen(path, "rb");
  if (!f) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_load: cannot open %s", path);
  CacheHeader h{};
  if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0)
  {
    std::fclose(f);
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_load: %s is not a grid cache file", path);
  }
  if (h.key != key || h.values != grid->values || h.n[0] != grid->n[0] || h.n[1] != grid->n[1] ||
      h.n[2] != grid->n[2] || h.rect != grid->rect || std::memcmp(h.cell, grid->cell, sizeof(h.cell)) != 0 ||
      h.count != grid->count)
  {
    std::fclose(f);
    return rg::fail(ctx, RG_ERR_INVALID,
                    "rg_grid_load: %s holds a different grid (key, scheme, dimensions or cell differ)", path);
  }
  std::vector<double> buf(std::min(kChunk, grid->count));
  for (size_t off = 0; off < grid->count; off += kChunk)
  {
    const size_t c = std::min(kChunk, grid->count - off);
    if (std::fread(buf.data(), sizeof(double), c, f) != c)
    {
      std::fclose(f);
      return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_load: %s is truncated", path);
    }
    hipError_t e = hipMemcpyAsync(grid->d + off, buf.data(), c * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
    {
      std::fclose(f);
      return rg::hipfail(ctx, e, "rg_grid_load");
    }
  }
  std::fclose(f);
  return RG_OK;
}

rg_status rg_grid_write_cube(rg_ctx* ctx, const rg_grid* grid, const char* path)
{
  if (!ctx || !grid || !path) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_write_cube: invalid arguments");
  RG_ENTER(ctx);
  FILE* f = std::fopen(path, "w");
  if (!f) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_write_cube: cannot open %s for writing", path);
  char num[64];
  auto put = [&](double v) {
    const std::to_chars_result r = std::to_chars(num, num + sizeof(num), v);
    std::fwrite(num, 1, size_t(r.ptr - num), f);
  };
  std::fputs("Cube interpolation file\nWritten by RASPA-3\n", f);
  // origin at zero; negative point counts mark lengths in Angstrom; voxel vectors = lattice vectors / point counts
  std::fputs("1 ", f);
  put(0.0);
  std::fputs(" ", f);
  put(0.0);
  std::fputs(" ", f);
  put(0.0);
  std::fputs("\n", f);
  for (int a = 0; a < 3; ++a)
  {
    std::fprintf(f, "%d", -grid->n[a]);
    for (int c = 0; c < 3; ++c)
    {
      std::fputs(" ", f);
      put(grid->cell[3 * a + c] / double(grid->n[a]));
    }
    std::fputs("\n", f);
  }
  std::fputs("1 0.0 0.0 0.0 0.0\n", f);
  // every stored value in memory order, in Kelvin, one per line (the reference writes the whole data array)
  const double e2k = rg::units().energyToKelvin;
  std::vector<double> buf(std::min(kChunk, grid->count));
  for (size_t off = 0; off < grid->count; off += kChunk)
  {
    const size_t c = std::min(kChunk, grid->count - off);
    hipError_t e = hipMemcpyAsync(buf.data(), grid->d + off, c * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
    {
      std::fclose(f);
      return rg::hipfail(ctx, e, "rg_grid_write_cube");
    }
    for (size_t i = 0; i < c; ++i)
    {
      put(buf[i] * e2k);
      std::fputc('\n', f);
    }
  }
  if (std::fclose(f) != 0) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_write_cube: writing %s failed", path);
  return RG_OK;
}

}  // extern "C"
