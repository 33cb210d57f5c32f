// io.cpp -- ESRI ASCII grids, polygon shapefiles, PLY and GeoTIFF writers.
#include "io.hpp"

#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace slabhost {

RasterData read_asc(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Failed to open ASC file: " + path);
    RasterData r;
    std::string key;
    // six header lines in the fixed order the reference reads them
    // (src/complex_terrain.cpp:39-44): ncols nrows xllcorner yllcorner cellsize nodata
    f >> key >> r.ncols >> key >> r.nrows >> key >> r.xllcorner >> key >> r.yllcorner >> key >> r.cellsize >> key >>
        r.nodata;
    if (!f || r.ncols < 1 || r.nrows < 1 || !(r.cellsize > 0)) throw std::runtime_error("Malformed ASC header: " + path);
    for (int i = 0; i < r.ncols; ++i) r.x.push_back(i * r.cellsize);
    for (int j = 0; j < r.nrows; ++j) r.y.push_back(j * r.cellsize);
    r.z.resize((size_t)r.ncols * r.nrows);
    for (int j = r.nrows - 1; j >= 0; --j)
        for (int i = 0; i < r.ncols; ++i)
            if (!(f >> r.z[(size_t)i * r.nrows + j])) throw std::runtime_error("ASC file truncated: " + path);
    return r;
}

void write_asc(const std::string& path, const RasterData& r) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("Cannot write " + path);
    f.precision(17);
    f << "ncols " << r.ncols << "\nnrows " << r.nrows << "\nxllcorner " << r.xllcorner << "\nyllcorner "
      << r.yllcorner << "\ncellsize " << r.cellsize << "\nNODATA_value " << r.nodata << "\n";
    for (int j = r.nrows - 1; j >= 0; --j) {
        for (int i = 0; i < r.ncols; ++i) f << (i ? " " : "") << r.z[(size_t)i * r.nrows + j];
        f << "\n";
    }
}

// --- shapefile (ESRI Shapefile Technical Description, 1998) -----------------

static int32_t be32(const unsigned char* p) {
    return (int32_t)((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3]);
}
static int32_t le32(const unsigned char* p) {
    int32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
static double le64(const unsigned char* p) {
    double v;
    std::memcpy(&v, p, 8);
    return v;
}

std::vector<std::pair<double, double>> read_shapefile_polygon(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Failed to open shapefile.");
    std::vector<unsigned char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (b.size() < 100 || be32(&b[0]) != 9994) throw std::runtime_error("Failed to open shapefile.");
    const int type = le32(&b[32]);
    if (type != 5 && type != 15 && type != 25) throw std::runtime_error("Not a polygon.");
    size_t off = 100;
    while (off + 8 <= b.size()) {
        const size_t len = (size_t)be32(&b[off + 4]) * 2;
        const size_t rec = off + 8;
        if (rec + len > b.size()) break;
        const int st = le32(&b[rec]);
        if (st == 0) { // null shape: next record
            off = rec + len;
            continue;
        }
        if (st != 5 && st != 15 && st != 25) throw std::runtime_error("Not a polygon.");
        const int nparts = le32(&b[rec + 36]), npoints = le32(&b[rec + 40]);
        if (nparts < 1 || npoints < 1) throw std::runtime_error("No exterior ring.");
        const size_t parts = rec + 44, pts = parts + 4 * (size_t)nparts;
        if (pts + 16 * (size_t)npoints > b.size()) throw std::runtime_error("Failed to get layer.");
        const int p0 = le32(&b[parts]);
        const int p1 = nparts > 1 ? le32(&b[parts + 4]) : npoints;
        std::vector<std::pair<double, double>> ring;
        for (int k = p0; k < p1; ++k) ring.emplace_back(le64(&b[pts + 16 * (size_t)k]), le64(&b[pts + 16 * (size_t)k + 8]));
        return ring;
    }
    throw std::runtime_error("No features found.");
}

static void put_be32(std::vector<unsigned char>& o, size_t at, int32_t v) {
    o[at] = (unsigned char)(v >> 24);
    o[at + 1] = (unsigned char)(v >> 16);
    o[at + 2] = (unsigned char)(v >> 8);
    o[at + 3] = (unsigned char)v;
}
template <typename T>
static void put_le(std::vector<unsigned char>& o, T v) {
    unsigned char buf[sizeof(T)];
    std::memcpy(buf, &v, sizeof(T));
    o.insert(o.end(), buf, buf + sizeof(T));
}

void write_shapefile_polygon(const std::string& path, const std::vector<std::pair<double, double>>& ring) {
    double xmin = 1e308, ymin = 1e308, xmax = -1e308, ymax = -1e308;
    for (auto& p : ring) {
        xmin = std::min(xmin, p.first);
        xmax = std::max(xmax, p.first);
        ymin = std::min(ymin, p.second);
        ymax = std::max(ymax, p.second);
    }
    std::vector<unsigned char> o(100, 0);
    put_be32(o, 0, 9994);
    std::vector<unsigned char> rec;
    put_le<int32_t>(rec, 5);
    for (double v : {xmin, ymin, xmax, ymax}) put_le<double>(rec, v);
    put_le<int32_t>(rec, 1);
    put_le<int32_t>(rec, (int32_t)ring.size());
    put_le<int32_t>(rec, 0);
    for (auto& p : ring) {
        put_le<double>(rec, p.first);
        put_le<double>(rec, p.second);
    }
    std::vector<unsigned char> hdr(8, 0);
    put_be32(hdr, 0, 1);
    put_be32(hdr, 4, (int32_t)(rec.size() / 2));
    o.insert(o.end(), hdr.begin(), hdr.end());
    o.insert(o.end(), rec.begin(), rec.end());
    put_be32(o, 24, (int32_t)(o.size() / 2));
    int32_t v = 1000;
    std::memcpy(&o[28], &v, 4);
    v = 5;
    std::memcpy(&o[32], &v, 4);
    const double bbox[8] = {xmin, ymin, xmax, ymax, 0, 0, 0, 0};
    std::memcpy(&o[36], bbox, sizeof(bbox));
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Cannot write " + path);
    f.write((const char*)o.data(), (std::streamsize)o.size());
}

// --- PLY --------------------------------------------------------------------

void write_ply(const std::string& path, int64_t np, const double* x, const double* y, const double* vx,
               const double* vy, const double* vz, const double* z, const double* h, const double* A) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Cannot write " + path);
    std::ostringstream hd;
    hd << "ply\nformat binary_little_endian 1.0\nelement vertex " << np << "\n";
    for (const char* n : {"x", "y", "vx", "vy", "vz", "z", "h", "A"}) hd << "property float " << n << "\n";
    hd << "end_header\n";
    const std::string s = hd.str();
    f.write(s.data(), (std::streamsize)s.size());
    std::vector<float> row((size_t)np * 8);
    for (int64_t p = 0; p < np; ++p) {
        float* r = &row[(size_t)p * 8];
        r[0] = (float)x[p];
        r[1] = (float)y[p];
        r[2] = (float)vx[p];
        r[3] = (float)vy[p];
        r[4] = (float)vz[p];
        r[5] = (float)z[p];
        r[6] = (float)h[p];
        r[7] = (float)A[p];
    }
    f.write((const char*)row.data(), (std::streamsize)(row.size() * 4));
}

// --- GeoTIFF (TIFF 6.0 + GeoTIFF 1.0 tags, uncompressed) --------------------

namespace {
struct Tag {
    uint16_t id, type;
    uint32_t count;
    std::vector<unsigned char> data;
};
template <typename T>
Tag tag(uint16_t id, uint16_t type, const std::vector<T>& v) {
    Tag t{id, type, (uint32_t)v.size(), {}};
    t.data.resize(v.size() * sizeof(T));
    std::memcpy(t.data.data(), v.data(), t.data.size());
    return t;
}
Tag ascii(uint16_t id, const std::string& s) {
    Tag t{id, 2, (uint32_t)(s.size() + 1), {}};
    t.data.assign(s.begin(), s.end());
    t.data.push_back(0);
    return t;
}
} // namespace

void write_geotiff(const std::string& path, int nx, int ny, double x_origin, double y_top, double dx,
                   const std::vector<std::string>& names, const std::vector<const float*>& bands, float nodata) {
    const int nb = (int)bands.size();
    const uint64_t row_bytes = (uint64_t)nx * nb * 4;
    const uint64_t img_bytes = row_bytes * (uint64_t)ny;
    if (img_bytes + 65536 > 0xFFFFFFF0ull) throw std::runtime_error("GeoTIFF larger than 4 GB: " + path);
    // rows of up to ~1 MB per strip
    const int rps = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ny, (1u << 20) / std::max<uint64_t>(row_bytes, 1)));
    const int nstrips = (ny + rps - 1) / rps;
    std::vector<Tag> tags;
    tags.push_back(tag<uint32_t>(256, 4, {(uint32_t)nx}));
    tags.push_back(tag<uint32_t>(257, 4, {(uint32_t)ny}));
    tags.push_back(tag<uint16_t>(258, 3, std::vector<uint16_t>(nb, 32)));
    tags.push_back(tag<uint16_t>(259, 3, {1}));
    tags.push_back(tag<uint16_t>(262, 3, {1}));
    tags.push_back(tag<uint32_t>(273, 4, std::vector<uint32_t>(nstrips, 0))); // patched below
    tags.push_back(tag<uint16_t>(277, 3, {(uint16_t)nb}));
    tags.push_back(tag<uint32_t>(278, 4, {(uint32_t)rps}));
    std::vector<uint32_t> counts(nstrips);
    for (int s = 0; s < nstrips; ++s) counts[s] = (uint32_t)(row_bytes * (uint64_t)std::min(rps, ny - s * rps));
    tags.push_back(tag<uint32_t>(279, 4, counts));
    tags.push_back(tag<uint16_t>(284, 3, {1}));
    if (nb > 1) tags.push_back(tag<uint16_t>(338, 3, std::vector<uint16_t>(nb - 1, 0)));
    tags.push_back(tag<uint16_t>(339, 3, std::vector<uint16_t>(nb, 3)));
    tags.push_back(tag<double>(33550, 12, {dx, dx, 0.0}));
    tags.push_back(tag<double>(33922, 12, {0.0, 0.0, 0.0, x_origin, y_top, 0.0}));
    // GeoKeyDirectory: version 1.1.0, one key GTRasterTypeGeoKey = RasterPixelIsArea
    tags.push_back(tag<uint16_t>(34735, 3, {1, 1, 0, 1, 1025, 0, 1, 1}));
    std::ostringstream md;
    md << "<GDALMetadata>\n";
    for (int b = 0; b < nb; ++b)
        md << "  <Item name=\"DESCRIPTION\" sample=\"" << b << "\" role=\"description\">" << names[b] << "</Item>\n";
    md << "</GDALMetadata>";
    tags.push_back(ascii(42112, md.str()));
    std::ostringstream nd;
    nd << nodata;
    tags.push_back(ascii(42113, nd.str()));

    // layout: header | IFD | out-of-line tag data | image strips
    const uint32_t ifd_off = 8;
    const uint32_t ifd_size = 2 + 12 * (uint32_t)tags.size() + 4;
    uint32_t data_off = ifd_off + ifd_size;
    std::vector<uint32_t> where(tags.size(), 0);
    for (size_t k = 0; k < tags.size(); ++k) {
        if (tags[k].data.size() > 4) {
            where[k] = data_off;
            data_off += (uint32_t)((tags[k].data.size() + 1) & ~(size_t)1);
        }
    }
    const uint32_t img_off = (data_off + 15) & ~15u;
    std::vector<uint32_t> offs(nstrips);
    for (int s = 0; s < nstrips; ++s) offs[s] = img_off + (uint32_t)(row_bytes * (uint64_t)s * rps);
    for (auto& t : tags)
        if (t.id == 273) std::memcpy(t.data.data(), offs.data(), offs.size() * 4);
    std::vector<unsigned char> head;
    head.insert(head.end(), {'I', 'I', 42, 0});
    put_le<uint32_t>(head, ifd_off);
    put_le<uint16_t>(head, (uint16_t)tags.size());
    for (size_t k = 0; k < tags.size(); ++k) {
        put_le<uint16_t>(head, tags[k].id);
        put_le<uint16_t>(head, tags[k].type);
        put_le<uint32_t>(head, tags[k].count);
        if (tags[k].data.size() > 4) {
            put_le<uint32_t>(head, where[k]);
        } else {
            unsigned char v[4] = {0, 0, 0, 0};
            std::memcpy(v, tags[k].data.data(), tags[k].data.size());
            head.insert(head.end(), v, v + 4);
        }
    }
    put_le<uint32_t>(head, 0);
    for (size_t k = 0; k < tags.size(); ++k) {
        if (tags[k].data.size() <= 4) continue;
        head.insert(head.end(), tags[k].data.begin(), tags[k].data.end());
        if (tags[k].data.size() & 1) head.push_back(0);
    }
    head.resize(img_off, 0);
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Could not create GeoTIFF: " + path);
    f.write((const char*)head.data(), (std::streamsize)head.size());
    std::vector<float> row((size_t)nx * nb);
    for (int r = 0; r < ny; ++r) {
        const int j = ny - 1 - r; // north-up
        for (int i = 0; i < nx; ++i)
            for (int b = 0; b < nb; ++b) row[(size_t)i * nb + b] = bands[b][(size_t)i * ny + j];
        f.write((const char*)row.data(), (std::streamsize)(row.size() * 4));
    }
    if (!f) throw std::runtime_error("Failed writing GeoTIFF: " + path);
}

} // namespace slabhost
