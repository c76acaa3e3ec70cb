// io.hip -- the reference's on-disk formats, host side.
//
//  * PETSc binary Vec files, the reference's output and restart format
//    (results/<field>%05d_0.dat, io.c:191-436, 657-745, 887-936, 1023-1072,
//    1924-2057; ParticleMotion.c:1168-1260).  The format belongs to PETSc
//    3.19.2 (the reference's 64-bit-index build): a big-endian PetscInt
//    VEC_FILE_CLASSID (1211214), a big-endian PetscInt length N, then N
//    big-endian doubles, with an optional "<file>.info" text file holding
//    "-vecload_block_size <bs>".  Files with 32-bit PetscInt headers (other
//    PETSc builds) are read as well.  DMDA vectors are stored in natural
//    order, which on one rank is the node-slot order [k][j][i][dof].
//  * The FDFGRID grid file (io.c:101-176, grid.c:356-383): a header line
//    "FDFGRID", the block count, one "IM JM KM" cell-count line per block,
//    then (IM+1)(JM+1)(KM+1) lines "x y z", k slowest and i fastest.
#include <errno.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "picgpu_internal.hpp"

using namespace picgpu;

namespace {

constexpr int64_t kVecClassId = 1211214;  // PETSc VEC_FILE_CLASSID
constexpr size_t kChunk = 1 << 20;        // doubles per I/O chunk

inline uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }
inline uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

struct File {
  FILE *f = nullptr;
  explicit File(FILE *p) : f(p) {}
  ~File() {
    if (f) fclose(f);
  }
};

int io_error(const char *what, const char *path) {
  set_error("%s: %s (%s)", what, path, strerror(errno));
  return PICGPU_E_ARG;
}

// Reads the header; *hdr receives its size in bytes (8 or 16).
int read_vec_header(FILE *f, const char *path, int64_t *len, int *hdr) {
  unsigned char b[16];
  if (fread(b, 1, 8, f) != 8) return io_error("short PETSc Vec file", path);
  uint32_t w0, w1;
  memcpy(&w0, b, 4);
  memcpy(&w1, b + 4, 4);
  if ((int64_t)bswap32(w0) == kVecClassId) {  // 32-bit PetscInt header
    *len = (int64_t)(int32_t)bswap32(w1);
    *hdr = 8;
  } else {
    uint64_t c;
    memcpy(&c, b, 8);
    if ((int64_t)bswap64(c) != kVecClassId) {
      set_error("%s is not a PETSc binary Vec file (no VEC_FILE_CLASSID)", path);
      return PICGPU_E_ARG;
    }
    if (fread(b + 8, 1, 8, f) != 8) return io_error("short PETSc Vec file", path);
    uint64_t n;
    memcpy(&n, b + 8, 8);
    *len = (int64_t)bswap64(n);
    *hdr = 16;
  }
  if (*len < 0) {
    set_error("%s: negative Vec length", path);
    return PICGPU_E_ARG;
  }
  return PICGPU_OK;
}

int write_vec(const char *path, int64_t len, int bs, const double *data) {
  File fp(fopen(path, "wb"));
  if (!fp.f) return io_error("cannot create", path);
  uint64_t h[2] = {bswap64((uint64_t)kVecClassId), bswap64((uint64_t)len)};
  if (fwrite(h, 8, 2, fp.f) != 2) return io_error("write failed", path);
  std::vector<uint64_t> buf(std::min<size_t>(kChunk, (size_t)len));
  for (int64_t s = 0; s < len; s += (int64_t)kChunk) {
    const size_t m = (size_t)std::min<int64_t>((int64_t)kChunk, len - s);
    memcpy(buf.data(), data + s, m * 8);
    for (size_t q = 0; q < m; ++q) buf[q] = bswap64(buf[q]);
    if (fwrite(buf.data(), 8, m, fp.f) != m) return io_error("write failed", path);
  }
  std::string info = std::string(path) + ".info";
  File fi(fopen(info.c_str(), "w"));
  if (!fi.f) return io_error("cannot create", info.c_str());
  if (bs > 1 && fprintf(fi.f, "-vecload_block_size %d\n", bs) < 0) return io_error("write failed", info.c_str());
  return PICGPU_OK;
}

int read_vec(const char *path, int64_t expected, double *out) {
  File fp(fopen(path, "rb"));
  if (!fp.f) {
    set_error("Restart file not found: %s", path);  // io.c:290-291
    return PICGPU_E_ARG;
  }
  int64_t len;
  int hdr;
  PG_CHECK(read_vec_header(fp.f, path, &len, &hdr));
  if (len != expected) {  // io.c:301-305
    set_error("File %s holds %" PRId64 " entries - expected %" PRId64, path, len, expected);
    return PICGPU_E_ARG;
  }
  std::vector<uint64_t> buf(std::min<size_t>(kChunk, (size_t)std::max<int64_t>(len, 1)));
  for (int64_t s = 0; s < len; s += (int64_t)kChunk) {
    const size_t m = (size_t)std::min<int64_t>((int64_t)kChunk, len - s);
    if (fread(buf.data(), 8, m, fp.f) != m) return io_error("short PETSc Vec file", path);
    for (size_t q = 0; q < m; ++q) buf[q] = bswap64(buf[q]);
    memcpy(out + s, buf.data(), m * 8);
  }
  return PICGPU_OK;
}

std::string field_path(const char *dir, const char *field, int64_t step, const char *ext) {
  char name[512];
  snprintf(name, sizeof(name), "%s%05" PRId64 "_0.%s", field, step, ext);  // io.c:680-682
  std::string d = dir && *dir ? dir : "results";
  return d + "/" + name;
}

bool file_exists(const std::string &p) {
  FILE *f = fopen(p.c_str(), "rb");
  if (!f) return false;
  fclose(f);
  return true;
}

}  // namespace

extern "C" {

int picgpu_field_file_name(char *buf, size_t size, const char *dir, const char *field, int64_t step,
                           const char *ext) {
  if (!buf || !field || !ext || step < 0) return PICGPU_E_ARG;
  const std::string p = field_path(dir, field, step, ext);
  if (p.size() + 1 > size) {
    set_error("file name buffer too small (%zu bytes needed)", p.size() + 1);
    return PICGPU_E_ARG;
  }
  memcpy(buf, p.c_str(), p.size() + 1);
  return PICGPU_OK;
}

int picgpu_petsc_vec_length(const char *path, int64_t *len) {
  if (!path || !len) return PICGPU_E_ARG;
  File fp(fopen(path, "rb"));
  if (!fp.f) {
    set_error("Restart file not found: %s", path);
    return PICGPU_E_ARG;
  }
  int hdr;
  return read_vec_header(fp.f, path, len, &hdr);
}

int picgpu_read_petsc_vec(const char *path, int64_t len, double *out) {
  if (!path || len < 0 || (len > 0 && !out)) return PICGPU_E_ARG;
  return read_vec(path, len, out);
}

int picgpu_write_petsc_vec(const char *path, int64_t len, int block_size, const double *data) {
  if (!path || len < 0 || (len > 0 && !data) || block_size < 1) return PICGPU_E_ARG;
  return write_vec(path, len, block_size, data);
}

// ReadGridFile (io.c:101-176) for the sizes of block `block`, then the
// coordinate read of AssignGridCoordinates (grid.c:356-383): skip nblk + 2
// lines, read (IM+1)(JM+1)(KM+1) "x y z" triples.  coor may be NULL (sizes
// only).  As in the reference, the skip assumes the "FDFGRID" header line.
int picgpu_read_grid_file(const char *path, int64_t *IM, int64_t *JM, int64_t *KM, int64_t *nblk, double *coor) {
  if (!path || !IM || !JM || !KM) return PICGPU_E_ARG;
  File fp(fopen(path, "r"));
  if (!fp.f) {
    set_error("Cannot open file: %s", path);
    return PICGPU_E_ARG;
  }
  char tok[32] = {0};
  if (fscanf(fp.f, "%31s", tok) != 1) {
    set_error("Empty grid file: %s", path);
    return PICGPU_E_ARG;
  }
  long long nb = 0;
  if (strcmp(tok, "FDFGRID") == 0) {
    if (fscanf(fp.f, "%lld", &nb) != 1) {
      set_error("Expected number of blocks after \"FDFGRID\" in %s", path);
      return PICGPU_E_ARG;
    }
  } else {
    nb = strtol(tok, nullptr, 10);
  }
  if (nb == 0) nb = 1;
  if (nb < 0) {
    set_error("negative block count in %s", path);
    return PICGPU_E_ARG;
  }
  long long im = 0, jm = 0, km = 0;
  for (long long b = 0; b < nb; ++b) {
    long long a, c, e;
    if (fscanf(fp.f, "%lld %lld %lld", &a, &c, &e) != 3) {
      set_error("Expected 3 integers for block %lld in %s", b, path);
      return PICGPU_E_ARG;
    }
    if (b == 0) {
      im = a;
      jm = c;
      km = e;
    }
  }
  if (im <= 0 || jm <= 0 || km <= 0) {
    set_error("Grid dimensions (IM, JM, KM) must be positive in %s", path);
    return PICGPU_E_ARG;
  }
  *IM = im;
  *JM = jm;
  *KM = km;
  if (nblk) *nblk = nb;
  if (!coor) return PICGPU_OK;
  rewind(fp.f);
  char line[256];
  for (long long s = 0; s < nb + 2; ++s)
    if (!fgets(line, sizeof(line), fp.f)) {
      set_error("Unexpected EOF while skipping grid header in %s", path);
      return PICGPU_E_ARG;
    }
  const int64_t nn = (im + 1) * (jm + 1) * (km + 1);
  for (int64_t q = 0; q < nn; ++q)
    if (fscanf(fp.f, "%le %le %le\n", &coor[3 * q], &coor[3 * q + 1], &coor[3 * q + 2]) != 3) {
      set_error("Error reading grid coordinates from file %s (node %" PRId64 ")", path, q);
      return PICGPU_E_ARG;
    }
  return PICGPU_OK;
}

// Writes a one-block FDFGRID file that picgpu_read_grid_file and the
// reference's reader accept; %.17g keeps every double exactly.
int picgpu_write_grid_file(const char *path, int64_t IM, int64_t JM, int64_t KM, const double *coor) {
  if (!path || !coor || IM <= 0 || JM <= 0 || KM <= 0) return PICGPU_E_ARG;
  File fp(fopen(path, "w"));
  if (!fp.f) return io_error("cannot create", path);
  if (fprintf(fp.f, "FDFGRID\n1\n%" PRId64 " %" PRId64 " %" PRId64 "\n", IM, JM, KM) < 0)
    return io_error("write failed", path);
  const int64_t nn = (IM + 1) * (JM + 1) * (KM + 1);
  for (int64_t q = 0; q < nn; ++q)
    if (fprintf(fp.f, "%.17g %.17g %.17g\n", coor[3 * q], coor[3 * q + 1], coor[3 * q + 2]) < 0)
      return io_error("write failed", path);
  return PICGPU_OK;
}

// WriteSwarmField("position"/"velocity") and WriteSimulationFields
// (simulation.c:304-306, io.c:887-936) for one output step: swarm fields in
// the device's particle order (block size 3), then vfield (Ucont), ufield
// (Ucat), pfield (the particle P average), nvfield (zero: no immersed body
// here) and ParticleCount, node-slot arrays.  Fields this context has not
// computed are written as zeros, as the reference's zero-initialised
// vectors would be.
int picgpu_write_output(picgpu_ctx *c, const char *dir, int64_t step) {
  if (!c || step < 0) return PICGPU_E_ARG;
  if (!c->have_grid) {
    set_error("no grid uploaded");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(settle_swarm(c));
  const int64_t n = c->n, nn = c->nnode;
  std::vector<double> buf((size_t)std::max<int64_t>(3 * n, 3 * nn));
  PG_CHECK(picgpu_download_particles(c, buf.data(), nullptr, nullptr, nullptr, nullptr, nullptr));
  PG_CHECK(write_vec(field_path(dir, "position", step, "dat").c_str(), 3 * n, 3, buf.data()));
  PG_CHECK(picgpu_download_particles(c, nullptr, buf.data(), nullptr, nullptr, nullptr, nullptr));
  PG_CHECK(write_vec(field_path(dir, "velocity", step, "dat").c_str(), 3 * n, 3, buf.data()));
  auto node_field = [&](int which, int64_t count, const char *name, int bs, bool available) -> int {
    if (available) {
      PG_CHECK(picgpu_download_grid_field(c, which, buf.data()));
    } else {
      memset(buf.data(), 0, sizeof(double) * (size_t)count);
    }
    return write_vec(field_path(dir, name, step, "dat").c_str(), count, bs, buf.data());
  };
  const bool eul = c->eul.have_metrics;
  PG_CHECK(node_field(PICGPU_FIELD_UCONT, 3 * nn, "vfield", 3, eul));
  if (eul) {
    PG_CHECK(node_field(PICGPU_FIELD_UCAT, 3 * nn, "ufield", 3, true));
  } else {
    // the particle loop's cell-linear copy, back in node slots
    memset(buf.data(), 0, sizeof(double) * 3 * (size_t)nn);
    if (c->have_ucat) {
      std::vector<unsigned char> h(c->rsize * 3 * (size_t)c->ncell);
      PG_HIP(hipMemcpy(h.data(), c->ucat.ptr, h.size(), hipMemcpyDeviceToHost));
      for (int64_t k = 0; k < c->KM; ++k)
        for (int64_t j = 0; j < c->JM; ++j)
          for (int64_t i = 0; i < c->IM; ++i) {
            const int64_t cc = (k * c->JM + j) * c->IM + i, nd = (k * (c->JM + 1) + j) * (c->IM + 1) + i;
            for (int q = 0; q < 3; ++q)
              buf[3 * nd + q] = c->rsize == 4 ? (double)((const float *)h.data())[3 * cc + q]
                                              : ((const double *)h.data())[3 * cc + q];
          }
    }
    PG_CHECK(write_vec(field_path(dir, "ufield", step, "dat").c_str(), 3 * nn, 3, buf.data()));
  }
  PG_CHECK(node_field(PICGPU_FIELD_P_AVERAGE, nn, "pfield", 1, c->have_pavg));
  PG_CHECK(node_field(PICGPU_FIELD_P_AVERAGE, nn, "nvfield", 1, false));
  PG_CHECK(node_field(PICGPU_FIELD_PARTICLE_COUNT, nn, "ParticleCount", 1, c->sorted || c->keyed));
  return PICGPU_OK;
}

// The restart read for step `step` (simulation.c:53-58 and
// ParticleSwarm.c:1187-1195): ufield -> Ucat and vfield -> Ucont (all five
// simulation files must exist and have the grid's size, io.c:437-488; P,
// Nvert and ParticleCount are checked and not kept), then the swarm from
// position (required; its length / 3 is the particle count,
// ParticleMotion.c:1168-1260) and velocity (if present, io.c:1974-2057),
// every particle UNINITIALIZED with no cell and PID 0..n-1
// (PrepareLoadedSwarmForRelocation, ParticleSwarm.c:1073-1130).
int picgpu_read_restart(picgpu_ctx *c, const char *dir, int64_t step) {
  if (!c || step < 0) return PICGPU_E_ARG;
  if (!c->have_grid) {
    set_error("no grid uploaded");
    return PICGPU_E_STATE;
  }
  const int64_t nn = c->nnode;
  std::vector<double> ucat((size_t)(3 * nn)), ucont((size_t)(3 * nn)), scratch((size_t)nn);
  PG_CHECK(read_vec(field_path(dir, "ufield", step, "dat").c_str(), 3 * nn, ucat.data()));
  PG_CHECK(read_vec(field_path(dir, "vfield", step, "dat").c_str(), 3 * nn, ucont.data()));
  PG_CHECK(read_vec(field_path(dir, "pfield", step, "dat").c_str(), nn, scratch.data()));
  PG_CHECK(read_vec(field_path(dir, "nvfield", step, "dat").c_str(), nn, scratch.data()));
  PG_CHECK(read_vec(field_path(dir, "ParticleCount", step, "dat").c_str(), nn, scratch.data()));
  PG_CHECK(picgpu_upload_ucat(c, ucat.data()));
  if (c->eul.have_metrics) PG_CHECK(picgpu_upload_ucont(c, ucont.data()));

  const std::string pp = field_path(dir, "position", step, "dat");
  int64_t len = 0;
  PG_CHECK(picgpu_petsc_vec_length(pp.c_str(), &len));
  if (len % 3 != 0) {
    set_error("%s holds %" PRId64 " entries, not a multiple of 3", pp.c_str(), len);
    return PICGPU_E_ARG;
  }
  const int64_t n = len / 3;
  std::vector<double> pos((size_t)len), vel((size_t)len, 0.0);
  PG_CHECK(read_vec(pp.c_str(), len, pos.data()));
  const std::string vp = field_path(dir, "velocity", step, "dat");
  if (file_exists(vp)) PG_CHECK(read_vec(vp.c_str(), len, vel.data()));
  std::vector<int64_t> cell((size_t)len, -1), status((size_t)n, PICGPU_UNINITIALIZED), pid((size_t)n);
  for (int64_t p = 0; p < n; ++p) pid[(size_t)p] = p;
  return picgpu_upload_particles(c, n, pos.data(), vel.data(), cell.data(), nullptr, status.data(), pid.data(),
                                 nullptr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// VTK XML output (io.c:1264-1815, postprocess.c:63-520): appended raw
// little-endian blocks, each led by its byte count as the reference's
// 64-bit PetscInt, offsets counted from the byte after "_".  The one
// difference from the reference's bytes: the VTKFile element also says
// header_type="UInt64", which is what the 8-byte counts need for VTK readers
// (without it they take 4-byte counts and misread every array).
// ---------------------------------------------------------------------------
namespace {

bool put_block(FILE *f, const void *data, int64_t count, size_t elem) {
  const int64_t bytes = count * (int64_t)elem;  // WriteVTKAppendedBlock, io.c:1264-1275
  if (fwrite(&bytes, sizeof(bytes), 1, f) != 1) return false;
  return count == 0 || fwrite(data, elem, (size_t)count, f) == (size_t)count;
}

int write_vts(const char *path, int64_t IM, int64_t JM, int64_t KM, const double *coor, const char *name, int ncomp,
              const double *field) {
  File fp(fopen(path, "wb"));
  if (!fp.f) return io_error("cannot create", path);
  const int64_t mx = IM + 1, my = JM + 1, mz = KM + 1, nn = mx * my * mz;
  int64_t off = 0;
  FILE *f = fp.f;
  fprintf(f, "<?xml version=\"1.0\"?>\n");
  fprintf(f, "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
  fprintf(f, "  <StructuredGrid WholeExtent=\"%d %" PRId64 " %d %" PRId64 " %d %" PRId64 "\">\n", 0, mx - 1, 0, my - 1,
          0, mz - 1);
  fprintf(f, "    <Piece Extent=\"%d %" PRId64 " %d %" PRId64 " %d %" PRId64 "\">\n", 0, mx - 1, 0, my - 1, 0, mz - 1);
  fprintf(f, "      <Points>\n");
  fprintf(f,
          "        <DataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" format=\"appended\" "
          "offset=\"%" PRId64 "\" />\n",
          off);
  off += 8 + 3 * nn * 8;
  fprintf(f, "      </Points>\n");
  fprintf(f, "      <PointData Scalars=\"%s\">\n", name);
  fprintf(f,
          "        <DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"appended\" "
          "offset=\"%" PRId64 "\" />\n",
          name, ncomp, off);
  fprintf(f, "      </PointData>\n");
  fprintf(f, "    </Piece>\n");
  fprintf(f, "  </StructuredGrid>\n");
  fprintf(f, "  <AppendedData encoding=\"raw\">\n");
  fprintf(f, "_");
  if (!put_block(f, coor, 3 * nn, sizeof(double)) || !put_block(f, field, (int64_t)ncomp * nn, sizeof(double)))
    return io_error("write failed", path);
  if (fprintf(f, "\n </AppendedData>\n</VTKFile>\n") < 0) return io_error("write failed", path);
  return PICGPU_OK;
}

int write_vtp(const char *path, int64_t n, const double *pos, const char *name, int ncomp, const double *field) {
  File fp(fopen(path, "wb"));
  if (!fp.f) return io_error("cannot create", path);
  FILE *f = fp.f;
  int64_t off = 0;
  fprintf(f, "<?xml version=\"1.0\"?>\n");
  fprintf(f, "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
  fprintf(f, "  <PolyData>\n");
  fprintf(f,
          "    <Piece NumberOfPoints=\"%" PRId64 "\" NumberOfVerts=\"%" PRId64
          "\" NumberOfLines=\"0\" NumberOfPolys=\"0\" NumberOfStrips=\"0\">\n",
          n, n);
  fprintf(f, "      <Points>\n");
  fprintf(f,
          "        <DataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" format=\"appended\" "
          "offset=\"%" PRId64 "\"/>\n",
          off);
  off += 8 + n * 3 * 8;
  fprintf(f, "      </Points>\n");
  fprintf(f, "      <PointData>\n");
  fprintf(f,
          "        <DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"appended\" "
          "offset=\"%" PRId64 "\"/>\n",
          name, ncomp, off);
  off += 8 + n * ncomp * 8;
  fprintf(f, "      </PointData>\n");
  fprintf(f, "      <Verts>\n");
  fprintf(f, "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"appended\" offset=\"%" PRId64 "\"/>\n",
          off);
  off += 8 + n * 8;
  fprintf(f, "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" offset=\"%" PRId64 "\"/>\n", off);
  fprintf(f, "      </Verts>\n");
  fprintf(f, "    </Piece>\n");
  fprintf(f, "  </PolyData>\n");
  fprintf(f, "  <AppendedData encoding=\"raw\">\n");
  fprintf(f, "_");
  // one vertex per particle: connectivity i, offsets i + 1 (postprocess.c:205-213)
  std::vector<int64_t> ids((size_t)n);
  if (!put_block(f, pos, 3 * n, sizeof(double)) || !put_block(f, field, (int64_t)ncomp * n, sizeof(double)))
    return io_error("write failed", path);
  for (int64_t i = 0; i < n; ++i) ids[(size_t)i] = i;
  if (!put_block(f, ids.data(), n, sizeof(int64_t))) return io_error("write failed", path);
  for (int64_t i = 0; i < n; ++i) ids[(size_t)i] = i + 1;
  if (!put_block(f, ids.data(), n, sizeof(int64_t))) return io_error("write failed", path);
  if (fprintf(f, "\n  </AppendedData>\n</VTKFile>\n") < 0) return io_error("write failed", path);
  return PICGPU_OK;
}

std::string vtk_path(const char *prefix, const char *name, int64_t step, const char *ext) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s_%05" PRId64 ".%s", name, step, ext);  // postprocess.c:296-297
  std::string d = prefix && *prefix ? prefix : ".";
  return d + "/" + buf;
}

}  // namespace

extern "C" {

int picgpu_write_vts(const char *path, int64_t IM, int64_t JM, int64_t KM, const double *coor, const char *name,
                     int ncomp, const double *field) {
  if (!path || !coor || !name || !field || IM <= 0 || JM <= 0 || KM <= 0 || (ncomp != 1 && ncomp != 3))
    return PICGPU_E_ARG;
  return write_vts(path, IM, JM, KM, coor, name, ncomp, field);
}

int picgpu_write_vtp(const char *path, int64_t n, const double *pos, const char *name, int ncomp,
                     const double *field) {
  if (!path || !name || n < 0 || (n > 0 && (!pos || !field)) || (ncomp != 1 && ncomp != 3)) return PICGPU_E_ARG;
  return write_vtp(path, n, pos, name, ncomp, field);
}

// WriteEulerianVTK + WriteParticleVTK (postprocess.c:451-520): Ucat, P,
// Ucont and ParticleCount on the grid nodes, the particles with their
// velocity.  Fields this context has not computed are written as zeros.
int picgpu_write_vtk(picgpu_ctx *c, const char *prefix, int64_t step) {
  if (!c || step < 0) return PICGPU_E_ARG;
  if (!c->have_grid) {
    set_error("no grid uploaded");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(settle_swarm(c));
  const int64_t nn = c->nnode, n = c->n;
  std::vector<double> coor((size_t)(3 * nn)), f((size_t)(3 * nn));
  PG_HIP(hipMemcpy(coor.data(), c->coor.ptr, sizeof(double) * coor.size(), hipMemcpyDeviceToHost));
  auto field = [&](int which, int64_t count, bool available) -> int {
    if (available) return picgpu_download_grid_field(c, which, f.data());
    memset(f.data(), 0, sizeof(double) * (size_t)count);
    return PICGPU_OK;
  };
  const bool eul = c->eul.have_metrics;
  PG_CHECK(field(PICGPU_FIELD_UCAT, 3 * nn, eul));
  if (!eul && c->have_ucat) {  // the particle loop's cell-linear copy, in node slots
    std::vector<unsigned char> h(c->rsize * 3 * (size_t)c->ncell);
    PG_HIP(hipMemcpy(h.data(), c->ucat.ptr, h.size(), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < c->KM; ++k)
      for (int64_t j = 0; j < c->JM; ++j)
        for (int64_t i = 0; i < c->IM; ++i) {
          const int64_t cc = (k * c->JM + j) * c->IM + i, nd = (k * (c->JM + 1) + j) * (c->IM + 1) + i;
          for (int q = 0; q < 3; ++q)
            f[3 * nd + q] = c->rsize == 4 ? (double)((const float *)h.data())[3 * cc + q]
                                          : ((const double *)h.data())[3 * cc + q];
        }
  }
  PG_CHECK(write_vts(vtk_path(prefix, "Ucat", step, "vts").c_str(), c->IM, c->JM, c->KM, coor.data(), "Ucat", 3,
                     f.data()));
  PG_CHECK(field(PICGPU_FIELD_P_AVERAGE, nn, c->have_pavg));
  PG_CHECK(write_vts(vtk_path(prefix, "P", step, "vts").c_str(), c->IM, c->JM, c->KM, coor.data(), "P", 1, f.data()));
  PG_CHECK(field(PICGPU_FIELD_UCONT, 3 * nn, eul));
  PG_CHECK(write_vts(vtk_path(prefix, "Ucont", step, "vts").c_str(), c->IM, c->JM, c->KM, coor.data(), "Ucont", 3,
                     f.data()));
  PG_CHECK(field(PICGPU_FIELD_PARTICLE_COUNT, nn, c->sorted || c->keyed));
  PG_CHECK(write_vts(vtk_path(prefix, "ParticleCount", step, "vts").c_str(), c->IM, c->JM, c->KM, coor.data(),
                     "ParticleCount", 1, f.data()));
  std::vector<double> pos((size_t)(3 * n)), vel((size_t)(3 * n));
  PG_CHECK(picgpu_download_particles(c, pos.data(), vel.data(), nullptr, nullptr, nullptr, nullptr));
  return write_vtp(vtk_path(prefix, "velocity", step, "vtp").c_str(), n, pos.data(), "velocity", 3, vel.data());
}

}  // extern "C"
