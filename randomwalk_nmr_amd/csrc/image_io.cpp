// image_io.cpp -- host-side image ingestion and the reference's text outputs:
// PNG slice stacks binarised as the reference does, the micro-CT config
// (uct.config), pore / interface statistics, ImageInfo.txt and walkers.csv.
//
// Reference behaviour restated (paths under src/backend/ of RandomWalk_NMR):
//   uct.config keys, "KEY: value" parsing   B/ConfigFiles/uct_config.cpp:14-115
//   slice file names (zero-padded id)       B/ConfigFiles/uct_config.cpp:165-196, uct_config.h:73-78
//   imread(path, IMREAD_COLOR) + binarise   B/NMR_Simulation/NMR_Simulation.cpp:484-617
//   pore count, porosity                    B/NMR_Simulation/NMR_Simulation.cpp:675-727, 919-923
//   pore-matrix interfaces, S/Vp            B/NMR_Simulation/NMR_Simulation.cpp:818-917
//   ImageInfo.txt, walkers.csv              B/NMR_Simulation/NMR_Simulation.cpp:1644-1712
//   internal field (uniform / file)         B/NMR_Simulation/InternalField.cpp:95-200
//
// The reference decodes slices with OpenCV's imread(path, 1): a colour
// (3-channel, 8-bit) image, alpha dropped, 16-bit samples reduced to their
// high byte, grey and palette expanded to colour; a pixel is solid when any
// colour channel is nonzero.  The decoder below (zlib inflate, the five PNG
// row filters, Adam7) applies the same rule to every PNG colour type and bit
// depth.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rwnmr/rwnmr_io.h"

namespace {

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

struct Png {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> palette;   // RGB triples
    std::vector<uint8_t> idat;
};

int channels_of(int ctype)
{
    switch (ctype) {
    case 0: return 1;
    case 2: return 3;
    case 3: return 1;
    case 4: return 2;
    case 6: return 4;
    default: return 0;
    }
}

bool depth_ok(int ctype, int depth)
{
    switch (ctype) {
    case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
    case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
    case 2: case 4: case 6: return depth == 8 || depth == 16;
    default: return false;
    }
}

rwnmr_status read_png_chunks(const char *path, Png &png, bool need_data)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        return RWNMR_ERR_INVALID;
    std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (buf.size() < 8 || memcmp(buf.data(), sig, 8) != 0)
        return RWNMR_ERR_INVALID;
    size_t pos = 8;
    bool have_ihdr = false;
    while (pos + 12 <= buf.size()) {
        const uint32_t len = be32(&buf[pos]);
        if (len > buf.size() - pos - 12)
            return RWNMR_ERR_INVALID;
        const uint8_t *type = &buf[pos + 4];
        const uint8_t *data = &buf[pos + 8];
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13)
                return RWNMR_ERR_INVALID;
            png.w = be32(data);
            png.h = be32(data + 4);
            png.depth = data[8];
            png.ctype = data[9];
            png.interlace = data[12];
            if (data[10] != 0 || data[11] != 0 || png.interlace > 1 || !depth_ok(png.ctype, png.depth) || png.w == 0 ||
                png.h == 0 || png.w > (1u << 24) || png.h > (1u << 24))
                return RWNMR_ERR_INVALID;
            have_ihdr = true;
            if (!need_data)
                return RWNMR_OK;
        } else if (!memcmp(type, "PLTE", 4)) {
            if (len % 3 != 0)
                return RWNMR_ERR_INVALID;
            png.palette.assign(data, data + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            png.idat.insert(png.idat.end(), data, data + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    return have_ihdr ? RWNMR_OK : RWNMR_ERR_INVALID;
}

uint8_t paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc)
        return (uint8_t)a;
    return (uint8_t)(pb <= pc ? b : c);
}

// Undo the row filters of one (sub)image of w x h pixels in place; rows are
// 1 filter byte + rowbytes.  Returns false on an unknown filter type.
bool unfilter(uint8_t *d, uint32_t w, uint32_t h, size_t rowbytes, size_t bpp)
{
    (void)w;
    std::vector<uint8_t> zero(rowbytes, 0);
    const uint8_t *prev = zero.data();
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t *row = d + (size_t)y * (rowbytes + 1);
        const int ft = row[0];
        uint8_t *x = row + 1;
        for (size_t i = 0; i < rowbytes; ++i) {
            const int a = i >= bpp ? x[i - bpp] : 0;
            const int b = prev[i];
            const int c = i >= bpp ? prev[i - bpp] : 0;
            switch (ft) {
            case 0: break;
            case 1: x[i] = (uint8_t)(x[i] + a); break;
            case 2: x[i] = (uint8_t)(x[i] + b); break;
            case 3: x[i] = (uint8_t)(x[i] + ((a + b) >> 1)); break;
            case 4: x[i] = (uint8_t)(x[i] + paeth(a, b, c)); break;
            default: return false;
            }
        }
        prev = x;
    }
    return true;
}

// Solid flag of pixel i of an unfiltered row (reference rule, see top).
uint8_t pixel_solid(const Png &png, const uint8_t *row, uint32_t i)
{
    const int d = png.depth;
    auto sample = [&](uint32_t s) -> uint32_t {   // s-th sample of the row
        if (d == 8)
            return row[s];
        if (d == 16)
            return (uint32_t)row[2 * s] << 8 | row[2 * s + 1];
        const uint32_t bit = s * (uint32_t)d;
        return (row[bit >> 3] >> (8 - d - (bit & 7))) & ((1u << d) - 1u);
    };
    auto nz8 = [&](uint32_t v) -> bool { return d == 16 ? (v >> 8) != 0 : v != 0; };
    switch (png.ctype) {
    case 0: return nz8(sample(i));
    case 2: return nz8(sample(3 * i)) || nz8(sample(3 * i + 1)) || nz8(sample(3 * i + 2));
    case 3: {
        const uint32_t idx = sample(i);
        if (3 * idx + 2 >= png.palette.size())
            return 1;   // index outside the palette: treated as solid
        return png.palette[3 * idx] || png.palette[3 * idx + 1] || png.palette[3 * idx + 2];
    }
    case 4: return nz8(sample(2 * i));
    case 6: return nz8(sample(4 * i)) || nz8(sample(4 * i + 1)) || nz8(sample(4 * i + 2));
    default: return 1;
    }
}

rwnmr_status decode_png(const char *path, uint32_t W, uint32_t H, uint8_t *solid)
{
    Png png;
    rwnmr_status s = read_png_chunks(path, png, true);
    if (s != RWNMR_OK)
        return s;
    if (png.w != W || png.h != H)
        return RWNMR_ERR_INVALID;
    if (png.ctype == 3 && png.palette.empty())
        return RWNMR_ERR_INVALID;
    const int ch = channels_of(png.ctype);
    const size_t bits_pp = (size_t)ch * png.depth;
    const size_t bpp = bits_pp >= 8 ? bits_pp / 8 : 1;
    // pass geometry (Adam7 or a single pass)
    static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
    const int passes = png.interlace ? 7 : 1;
    size_t total = 0;
    uint32_t pw[7], ph[7];
    for (int p = 0; p < passes; ++p) {
        if (png.interlace) {
            pw[p] = png.w > (uint32_t)ax0[p] ? (png.w - ax0[p] + adx[p] - 1) / adx[p] : 0;
            ph[p] = png.h > (uint32_t)ay0[p] ? (png.h - ay0[p] + ady[p] - 1) / ady[p] : 0;
        } else {
            pw[p] = png.w;
            ph[p] = png.h;
        }
        if (pw[p] && ph[p])
            total += (size_t)ph[p] * (1 + (pw[p] * bits_pp + 7) / 8);
    }
    std::vector<uint8_t> raw(total);
    uLongf out_len = (uLongf)total;
    if (uncompress(raw.data(), &out_len, png.idat.data(), (uLong)png.idat.size()) != Z_OK || out_len != total)
        return RWNMR_ERR_INVALID;
    size_t off = 0;
    for (int p = 0; p < passes; ++p) {
        if (!pw[p] || !ph[p])
            continue;
        const size_t rowbytes = (pw[p] * bits_pp + 7) / 8;
        if (!unfilter(raw.data() + off, pw[p], ph[p], rowbytes, bpp))
            return RWNMR_ERR_INVALID;
        for (uint32_t y = 0; y < ph[p]; ++y) {
            const uint8_t *row = raw.data() + off + (size_t)y * (rowbytes + 1) + 1;
            const uint32_t Y = png.interlace ? ay0[p] + y * ady[p] : y;
            for (uint32_t x = 0; x < pw[p]; ++x) {
                const uint32_t X = png.interlace ? ax0[p] + x * adx[p] : x;
                solid[(size_t)Y * W + X] = pixel_solid(png, row, x);
            }
        }
        off += (size_t)ph[p] * (rowbytes + 1);
    }
    return RWNMR_OK;
}

void copy_str(char *dst, size_t cap, const std::string &s)
{
    const size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
    memcpy(dst, s.data(), n);
    dst[n] = 0;
}

bool parse_int(const std::string &s, int32_t &v)
{
    try {
        v = (int32_t)std::stoi(s);
        return true;
    } catch (...) {
        return false;
    }
}

bool parse_double(const std::string &s, double &v)
{
    try {
        v = std::stod(s);
        return true;
    } catch (...) {
        return false;
    }
}

bool solid_at(const uint8_t *solid, int32_t W, int32_t H, int32_t x, int32_t y, int32_t z)
{
    return solid[((size_t)z * H + y) * W + x] != 0;
}

}  // namespace

extern "C" {

rwnmr_status rwnmr_png_size(const char *path, int32_t *width, int32_t *height)
{
    if (!path || !width || !height)
        return RWNMR_ERR_INVALID;
    Png png;
    rwnmr_status s = read_png_chunks(path, png, false);
    if (s != RWNMR_OK)
        return s;
    *width = (int32_t)png.w;
    *height = (int32_t)png.h;
    return RWNMR_OK;
}

rwnmr_status rwnmr_png_binarize(const char *path, int32_t width, int32_t height, uint8_t *solid)
{
    if (!path || !solid || width <= 0 || height <= 0)
        return RWNMR_ERR_INVALID;
    return decode_png(path, (uint32_t)width, (uint32_t)height, solid);
}

void rwnmr_uct_config_defaults(rwnmr_uct_config *cfg)
{
    if (!cfg)
        return;
    memset(cfg, 0, sizeof(*cfg));
    copy_str(cfg->dir_path, sizeof(cfg->dir_path), "path_to_folder_with_images");
    copy_str(cfg->filename, sizeof(cfg->filename), "imagefile_prefix");
    cfg->first_idx = 0;
    cfg->digits = 1;
    copy_str(cfg->extension, sizeof(cfg->extension), ".png");
    cfg->slices = 1;
    cfg->resolution = 1.0;
    cfg->voxel_division = 0;
    copy_str(cfg->img_files_list, sizeof(cfg->img_files_list), "Empty");
}

rwnmr_status rwnmr_uct_config_read(const char *path, rwnmr_uct_config *cfg)
{
    if (!path || !cfg)
        return RWNMR_ERR_INVALID;
    std::ifstream f(path);
    if (!f)
        return RWNMR_ERR_INVALID;
    std::string line;
    bool ok = true;
    while (std::getline(f, line)) {
        // every "token: " prefix of the line is a key candidate, the rest of
        // the line its value (uct_config.cpp:63-87)
        std::string s = line;
        size_t pos;
        while ((pos = s.find(": ")) != std::string::npos) {
            const std::string token = s.substr(0, pos), content = s.substr(pos + 2);
            s.erase(0, pos + 2);
            if (token == "DIR_PATH")
                copy_str(cfg->dir_path, sizeof(cfg->dir_path), content);
            else if (token == "FILENAME")
                copy_str(cfg->filename, sizeof(cfg->filename), content);
            else if (token == "FIRST_IDX")
                ok &= parse_int(content, cfg->first_idx);
            else if (token == "DIGITS")
                ok &= parse_int(content, cfg->digits);
            else if (token == "EXTENSION")
                copy_str(cfg->extension, sizeof(cfg->extension), content);
            else if (token == "SLICES")
                ok &= parse_int(content, cfg->slices);
            else if (token == "RESOLUTION")
                ok &= parse_double(content, cfg->resolution);
            else if (token == "VOXEL_DIVISION")
                ok &= parse_int(content, cfg->voxel_division);
            else if (token == "IMG_FILES_LIST")
                copy_str(cfg->img_files_list, sizeof(cfg->img_files_list), content);
        }
    }
    return ok ? RWNMR_OK : RWNMR_ERR_INVALID;
}

rwnmr_status rwnmr_uct_slice_path(const rwnmr_uct_config *cfg, int32_t slice, char *buf, uint32_t buflen)
{
    if (!cfg || !buf || buflen == 0 || slice < 0 || slice >= cfg->slices)
        return RWNMR_ERR_INVALID;
    std::string path;
    if (cfg->img_files_list[0] && strcmp(cfg->img_files_list, "Empty") != 0) {
        // explicit list: line `slice` of the file (uct_config.cpp:130-158)
        std::ifstream f(cfg->img_files_list);
        if (!f)
            return RWNMR_ERR_INVALID;
        std::string line;
        for (int32_t i = 0; i <= slice; ++i)
            if (!std::getline(f, line))
                return RWNMR_ERR_INVALID;
        path = line;
    } else {
        std::ostringstream id;
        id << std::setfill('0') << std::setw(cfg->digits < 0 ? 0 : cfg->digits) << (uint32_t)(cfg->first_idx + slice);
        path = std::string(cfg->dir_path) + cfg->filename + id.str() + cfg->extension;
    }
    if (path.size() + 1 > buflen)
        return RWNMR_ERR_INVALID;
    memcpy(buf, path.c_str(), path.size() + 1);
    return RWNMR_OK;
}

rwnmr_status rwnmr_uct_dims(const rwnmr_uct_config *cfg, int32_t *width, int32_t *height, int32_t *depth)
{
    if (!cfg || !width || !height || !depth || cfg->slices <= 0)
        return RWNMR_ERR_INVALID;
    char path[4096];
    rwnmr_status s = rwnmr_uct_slice_path(cfg, 0, path, sizeof(path));
    if (s != RWNMR_OK)
        return s;
    if ((s = rwnmr_png_size(path, width, height)) != RWNMR_OK)
        return s;
    *depth = cfg->slices;
    return RWNMR_OK;
}

rwnmr_status rwnmr_uct_load(const rwnmr_uct_config *cfg, int32_t width, int32_t height, int32_t depth, uint8_t *solid)
{
    if (!cfg || !solid || width <= 0 || height <= 0 || depth != cfg->slices)
        return RWNMR_ERR_INVALID;
    char path[4096];
    for (int32_t z = 0; z < depth; ++z) {
        rwnmr_status s = rwnmr_uct_slice_path(cfg, z, path, sizeof(path));
        if (s != RWNMR_OK)
            return s;
        if ((s = rwnmr_png_binarize(path, width, height, solid + (size_t)z * width * height)) != RWNMR_OK)
            return s;
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_image_stats(const uint8_t *solid, int32_t width, int32_t height, int32_t depth, double voxel,
                               uint64_t *pores, double *porosity, uint64_t *interfaces, double *svp)
{
    if (!solid || width <= 0 || height <= 0 || depth <= 0)
        return RWNMR_ERR_INVALID;
    uint64_t np = 0, ni = 0;
    const bool dim3 = depth > 1;
    for (int32_t z = 0; z < depth; ++z) {
        const int32_t nz = (z + 1) % depth;
        for (int32_t y = 0; y < height; ++y) {
            const int32_t ny = (y + 1) % height;
            for (int32_t x = 0; x < width; ++x) {
                const int32_t nx = (x + 1) % width;
                const bool c = solid_at(solid, width, height, x, y, z);
                np += c ? 0u : 1u;
                ni += c != solid_at(solid, width, height, nx, y, z);
                ni += c != solid_at(solid, width, height, x, ny, z);
                if (dim3)
                    ni += c != solid_at(solid, width, height, x, y, nz);
            }
        }
    }
    const double vol = (double)width * height * depth;
    if (pores)
        *pores = np;
    if (porosity)
        *porosity = (double)np / vol;
    if (interfaces)
        *interfaces = ni;
    if (svp)
        *svp = ((double)ni * voxel * voxel) / ((double)np * voxel * voxel * voxel);
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_image_info(const char *path, const char *image_path, int32_t width, int32_t height,
                                    int32_t depth, double resolution, int32_t voxel_shift, double step_length,
                                    double porosity, double svp)
{
    if (!path)
        return RWNMR_ERR_INVALID;
    std::ofstream f(path);
    if (!f)
        return RWNMR_ERR_INVALID;
    f << "image path: " << (image_path ? image_path : "") << std::endl;
    f << "width(x): " << width << std::endl;
    f << "height(y): " << height << std::endl;
    f << "depth(z): " << depth << std::endl;
    f << "resolution: " << resolution << std::endl;
    f << "voxel shift: " << voxel_shift << std::endl;
    f << "step length: " << step_length << std::endl;
    f << "porosity: " << porosity << std::endl;
    f << "SVp: " << svp << std::endl;
    return f ? RWNMR_OK : RWNMR_ERR_INVALID;
}

rwnmr_status rwnmr_write_walkers_csv(const char *path, uint64_t n, const int32_t *x0, const int32_t *y0,
                                     const int32_t *z0, const int32_t *x, const int32_t *y, const int32_t *z,
                                     const uint32_t *collisions, const double *xi_rate, const double *energy,
                                     const uint64_t *seed)
{
    if (!path || (n && (!x0 || !y0 || !z0 || !x || !y || !z || !collisions || !xi_rate || !energy || !seed)))
        return RWNMR_ERR_INVALID;
    std::ofstream f(path);
    if (!f)
        return RWNMR_ERR_INVALID;
    f << "PositionXi,PositionYi,PositionZi,PositionXf,PositionYf,PositionZf,Collisions,XIRate,Energy,RNGSeed"
      << std::endl;
    f << std::setprecision(6);
    for (uint64_t i = 0; i < n; ++i)
        f << x0[i] << "," << y0[i] << "," << z0[i] << "," << x[i] << "," << y[i] << "," << z[i] << "," << collisions[i]
          << "," << xi_rate[i] << "," << energy[i] << "," << seed[i] << std::endl;
    return f ? RWNMR_OK : RWNMR_ERR_INVALID;
}

rwnmr_status rwnmr_internal_field_uniform(const uint8_t *solid, int32_t width, int32_t height, int32_t depth,
                                          double resolution, double gradient, int32_t direction, double *field)
{
    if (!solid || !field || width <= 0 || height <= 0 || depth <= 0)
        return RWNMR_ERR_INVALID;
    const double step = gradient * 1.0e-6 * resolution;
    for (int32_t z = 0; z < depth; ++z)
        for (int32_t y = 0; y < height; ++y)
            for (int32_t x = 0; x < width; ++x) {
                const size_t i = ((size_t)z * height + y) * width + x;
                const double v = direction == 0 ? x * step : direction == 1 ? y * step : direction == 2 ? z * step : 0.0;
                field[i] = solid[i] ? 0.0 : v;
            }
    return RWNMR_OK;
}

rwnmr_status rwnmr_internal_field_read(const char *path, int32_t width, int32_t height, int32_t depth, double *field)
{
    if (!path || !field || width <= 0 || height <= 0 || depth <= 0)
        return RWNMR_ERR_INVALID;
    std::ifstream f(path, std::ios::binary);
    if (!f)
        return RWNMR_ERR_INVALID;
    f.seekg(0, std::ios::end);
    const std::streamoff bytes = f.tellg();
    const size_t n = (size_t)width * height * depth;
    if (bytes < 0 || (size_t)bytes != n * sizeof(double))
        return RWNMR_ERR_INVALID;
    f.seekg(0, std::ios::beg);
    f.read(reinterpret_cast<char *>(field), (std::streamsize)(n * sizeof(double)));
    return f ? RWNMR_OK : RWNMR_ERR_INVALID;
}

}  // extern "C"
