// io.hpp -- host-side input/output of the front end: ESRI ASCII grids,
// polygon shapefiles, binary PLY particle frames and GeoTIFF rasters.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace slabhost {

// ESRI ASCII grid as src/complex_terrain.cpp:31-65 reads it: x_i = i*cellsize,
// y_j = j*cellsize, rows reversed so z[i*nrows + j] has y up.
struct RasterData {
    std::vector<double> x, y, z;
    int ncols = 0, nrows = 0;
    double xllcorner = 0, yllcorner = 0, cellsize = 0, nodata = 0;
};
RasterData read_asc(const std::string& path);
void write_asc(const std::string& path, const RasterData& r);

// First polygon record's outer ring of an ESRI shapefile (shape type 5 /
// 15 / 25), as src/polygon_tools.cpp:23-65 reads it through OGR.
std::vector<std::pair<double, double>> read_shapefile_polygon(const std::string& path);
void write_shapefile_polygon(const std::string& path, const std::vector<std::pair<double, double>>& ring);

// Binary little-endian PLY, element "vertex" with float properties
// x y vx vy vz z h A (src/save_function.cpp:36-76 through tinyply).
void write_ply(const std::string& path, int64_t np, const double* x, const double* y, const double* vx,
               const double* vy, const double* vz, const double* z, const double* h, const double* A);

// GeoTIFF, Float32, bands interleaved per pixel, north-up rows, geotransform
// (x_origin, dx, 0, y_top, 0, -dx), GDAL nodata and band descriptions.
// bands[b][i*ny + j] in the node order of the grid (i along x, j along y).
void write_geotiff(const std::string& path, int nx, int ny, double x_origin, double y_top, double dx,
                   const std::vector<std::string>& names, const std::vector<const float*>& bands, float nodata);

} // namespace slabhost
