// ParaView export of the particle store.  The reference exports its agents every step through
// BioDynaMo's visualization (bdm.toml:6-13: export = true, interval = 1, ChargedAntiparticle with the
// additional data member species_id, charged_antiparticle.h:73) next to the agents' position and
// diameter (1e-6 m, constants.h:38, set in charged_antiparticle.cc:9).  BioDynaMo is not part of this
// library, so the same content is written as VTK XML: one UnstructuredGrid (.vtu) of vertex cells per
// exported step, and a .pvd collection that ParaView opens as one time series.
#include <cstdio>
#include <string>
#include <vector>

#include "antimatter_amd/am_params.h"

namespace am {
int Fail(int code, const std::string& msg);
}  // namespace am

namespace {

constexpr double kDefaultParticleDiameter = 1e-6;  // m

}  // namespace

extern "C" int am_export_vtu(am_store* s, const char* path) {
  if (!s || !path) return am::Fail(AM_ERR_INVALID, "am_export_vtu: NULL argument");
  int64_t n = 0, n_active = 0;
  int rc = am_store_count(s, &n, &n_active);
  if (rc) return rc;
  std::vector<double> pos(3 * (size_t)n);
  std::vector<uint8_t> gone(n), species(n);
  if (n > 0) {
    if ((rc = am_store_download(s, pos.data(), nullptr, nullptr, gone.data(), nullptr))) return rc;
    if ((rc = am_store_download_species(s, species.data()))) return rc;
  }
  std::vector<int64_t> keep;  // agents still in the simulation (annihilated ones are removed)
  keep.reserve(n);
  for (int64_t i = 0; i < n; ++i)
    if (!gone[i]) keep.push_back(i);
  const size_t m = keep.size();
  FILE* f = std::fopen(path, "w");
  if (!f) return am::Fail(AM_ERR_INVALID, std::string("am_export_vtu: cannot write ") + path);
  std::fprintf(f,
               "<?xml version=\"1.0\"?>\n"
               "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
               "<UnstructuredGrid>\n<Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">\n",
               m, m);
  std::fprintf(f, "<PointData Scalars=\"species_id\">\n<DataArray type=\"Int32\" Name=\"species_id\" format=\"ascii\">\n");
  for (int64_t i : keep) std::fprintf(f, "%d\n", (int)species[i]);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"Float64\" Name=\"diameter_\" format=\"ascii\">\n");
  for (size_t k = 0; k < m; ++k) std::fprintf(f, "%.17g\n", kDefaultParticleDiameter);
  std::fprintf(f, "</DataArray>\n</PointData>\n<Points>\n"
                  "<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int64_t i : keep)
    std::fprintf(f, "%.17g %.17g %.17g\n", pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  std::fprintf(f, "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
  for (size_t k = 0; k < m; ++k) std::fprintf(f, "%zu\n", k);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  for (size_t k = 0; k < m; ++k) std::fprintf(f, "%zu\n", k + 1);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (size_t k = 0; k < m; ++k) std::fprintf(f, "1\n");  // VTK_VERTEX
  std::fprintf(f, "</DataArray>\n</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) return am::Fail(AM_ERR_INVALID, std::string("am_export_vtu: write failed: ") + path);
  return AM_OK;
}

extern "C" int am_export_pvd(const char* path, const char* const* files, const double* times, int64_t count) {
  if (!path || (count > 0 && (!files || !times)) || count < 0)
    return am::Fail(AM_ERR_INVALID, "am_export_pvd: invalid argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return am::Fail(AM_ERR_INVALID, std::string("am_export_pvd: cannot write ") + path);
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\">\n<Collection>\n");
  for (int64_t k = 0; k < count; ++k)
    std::fprintf(f, "<DataSet timestep=\"%.17g\" part=\"0\" file=\"%s\"/>\n", times[k], files[k]);
  std::fprintf(f, "</Collection>\n</VTKFile>\n");
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) return am::Fail(AM_ERR_INVALID, std::string("am_export_pvd: write failed: ") + path);
  return AM_OK;
}
