// On-disk formats of spade::io over the C ABI: the parallel VTK
// StructuredGrid output (io/io_vtk.h:131-288) and the raw per-block array
// image (io/io_native.h:19-55).  Host code only; device arrays reach the host
// through the grid's staging buffer, a chunk of blocks at a time.
#include "api_internal.hpp"
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <sstream>

using namespace spd;

namespace {

const char* const kPrimNames[5] = {"P", "T", "U", "V", "W"};

// RFC 4648 base64 of n bytes, '='-padded to a multiple of four characters.
void put_base64(std::ostream& os, const unsigned char* p, size_t n)
{
    static const char A[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve(4*((n + 2)/3));
    size_t i = 0;
    for (; i + 3 <= n; i += 3)
    {
        const uint32_t v = (uint32_t(p[i]) << 16) | (uint32_t(p[i + 1]) << 8) | uint32_t(p[i + 2]);
        out += A[v >> 18]; out += A[(v >> 12) & 63]; out += A[(v >> 6) & 63]; out += A[v & 63];
    }
    if (i < n)
    {
        const bool two = (i + 1 < n);
        const uint32_t v = (uint32_t(p[i]) << 16) | (two ? uint32_t(p[i + 1]) << 8 : 0u);
        out += A[v >> 18]; out += A[(v >> 12) & 63];
        out += two ? A[(v >> 6) & 63] : '=';
        out += '=';
    }
    os << out;
}

// Inline binary payload of a DataArray as the reference streams it
// (core/base_64.h:63-69): the byte count as a 4-byte unsigned int and the
// data, each encoded and padded on its own.
void put_payload(std::ostream& os, const std::vector<double>& d)
{
    const uint32_t bytes = uint32_t(d.size()*sizeof(double));
    put_base64(os, reinterpret_cast<const unsigned char*>(&bytes), sizeof(bytes));
    put_base64(os, reinterpret_cast<const unsigned char*>(d.data()), d.size()*sizeof(double));
}

bool little_endian()
{
    const uint16_t one = 1;
    return *reinterpret_cast<const unsigned char*>(&one) == 1;
}

std::string endian_str() { return little_endian() ? "LittleEndian" : "BigEndian"; }

std::string block_file(const std::string& basename, long long lb)
{
    char num[32];
    std::snprintf(num, sizeof(num), "b%09lld.vts", lb);
    return (std::filesystem::path("data_" + basename) / num).string();
}

std::string joined_names(const char* const* names)
{
    std::string s;
    for (int v = 0; v < 5; ++v) { if (v) s += ","; s += names[v]; }
    return s;
}

void ensure_dir(const std::filesystem::path& p)
{
    std::error_code ec;
    if (std::filesystem::is_directory(p, ec)) return;
    std::filesystem::create_directories(p, ec);
    if (ec) throw IoError{"cannot create directory " + p.string() + ": " + ec.message()};
}

void check_stream(const std::ofstream& f, const std::string& path)
{
    if (!f) throw IoError{"cannot write " + path};
}

// <basename>.pvts (io_vtk.h:131-174): the pieces stack the blocks along z
void write_base_file(const spd_grid_desc& d, const std::string& path, const std::string& basename,
                     const char* const* names)
{
    const int* nc = d.cells_in_block;
    const long long nblk = (long long)d.num_blocks[0]*d.num_blocks[1]*d.num_blocks[2];
    std::ofstream f(path);
    if (!f) throw IoError{"cannot open " + path + " for writing"};
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"PStructuredGrid\" version=\"0.1\" byte_order=\"" << endian_str() << "\">\n";
    f << "<PStructuredGrid WholeExtent=\"0 " << nc[0] << " 0 " << nc[1] << " 0 " << (long long)nc[2]*nblk
      << "\" GhostLevel=\"" << std::min(d.num_exch, 0) << "\">\n";
    f << "<PCellData Scalars = \"" << joined_names(names) << "\">\n";
    for (int v = 0; v < 5; ++v) f << "<PDataArray type=\"Float64\" Name=\"" << names[v] << "\"/>\n";
    f << "</PCellData>\n";
    f << "<PPoints>\n";
    f << "<PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n";
    f << "</PPoints>\n";
    for (long long lb = 0; lb < nblk; ++lb)
        f << "<Piece Extent=\"0 " << nc[0] << " 0 " << nc[1] << " " << lb*nc[2] << " " << (lb + 1)*nc[2]
          << "\" Source=\"" << block_file(basename, lb) << "\"/>\n";
    f << "</PStructuredGrid>\n";
    f << "</VTKFile>\n";
    check_stream(f, path);
}

// data_<basename>/b<lb>.vts (io_vtk.h:176-273) from one block's image in the
// reference layout (AoS, variable fastest, ghosts stored)
void write_block_file(const spd_grid_desc& d, long long lb, const double* img, const std::string& path,
                      const char* const* names)
{
    const int* nc = d.cells_in_block;
    const int ng = d.num_exch;
    const int Nx = nc[0] + 2*ng, Ny = nc[1] + 2*ng;
    std::ofstream f(path);
    if (!f) throw IoError{"cannot open " + path + " for writing"};
    const std::string ext = "0 " + std::to_string(nc[0]) + " 0 " + std::to_string(nc[1]) + " 0 " + std::to_string(nc[2]);
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"" << endian_str() << "\">\n";
    f << "<StructuredGrid WholeExtent=\"" << ext << "\">\n";
    f << "<Piece Extent=\"" << ext << "\">\n";
    f << "<CellData Scalars=\"" << joined_names(names) << "\">\n";
    std::vector<double> data((size_t)nc[0]*nc[1]*nc[2]);
    for (int v = 0; v < 5; ++v)
    {
        size_t o = 0;
        for (int k = 0; k < nc[2]; ++k) for (int j = 0; j < nc[1]; ++j) for (int i = 0; i < nc[0]; ++i)
            data[o++] = img[5*((size_t)(i + ng) + (size_t)Nx*((j + ng) + (size_t)Ny*(k + ng))) + v];
        f << "<DataArray type=\"Float64\" Name=\"" << names[v] << "\" format=\"binary\">\n";
        put_payload(f, data);
        f << "\n</DataArray>\n";
    }
    f << "</CellData>\n";
    f << "<Points>\n";
    f << "<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"binary\">\n";
    // node coordinates x = map(box min + i dx) (grid/grid_geometry.h:51-69)
    double mn[3], dx[3];
    for (int a = 0; a < 3; ++a) block_box(d, (int)lb, a, mn[a], dx[a]);
    std::vector<double> pts;
    pts.reserve(3*(size_t)(nc[0] + 1)*(nc[1] + 1)*(nc[2] + 1));
    for (int k = 0; k <= nc[2]; ++k) for (int j = 0; j <= nc[1]; ++j) for (int i = 0; i <= nc[0]; ++i)
    {
        pts.push_back(map_coord(d.map[0], mn[0] + i*dx[0]));
        pts.push_back(map_coord(d.map[1], mn[1] + j*dx[1]));
        pts.push_back(map_coord(d.map[2], mn[2] + k*dx[2]));
    }
    put_payload(f, pts);
    f << "\n</DataArray>\n";
    f << "</Points>\n";
    f << "</Piece>\n";
    f << "</StructuredGrid>\n";
    f << "</VTKFile>\n";
    check_stream(f, path);
}

// Writes the base file and the block files; fetch(lb0, nb, buf) fills buf
// with the images of blocks [lb0, lb0 + nb).
template <typename Fetch>
void output_vtk(const spd_grid_desc& d, const char* out_dir, const char* basename, const char* const* names,
                int chunk_blocks, Fetch&& fetch)
{
    require(out_dir && basename && *basename, "output directory and base name are required");
    if (!names) names = kPrimNames;
    for (int v = 0; v < 5; ++v) require(names[v] && *names[v], "variable names must be non-empty");
    const std::filesystem::path out(out_dir);
    ensure_dir(out);
    ensure_dir(out / ("data_" + std::string(basename)));
    write_base_file(d, (out / (std::string(basename) + ".pvts")).string(), basename, names);
    const int nblk = d.num_blocks[0]*d.num_blocks[1]*d.num_blocks[2];
    const size_t blk = 5*(size_t)(d.cells_in_block[0] + 2*d.num_exch)*(d.cells_in_block[1] + 2*d.num_exch)*
                       (d.cells_in_block[2] + 2*d.num_exch);
    std::vector<double> buf;
    for (int lb0 = 0; lb0 < nblk; lb0 += chunk_blocks)
    {
        const int nb = std::min(chunk_blocks, nblk - lb0);
        const double* src = fetch(lb0, nb, buf);
        for (int b = 0; b < nb; ++b)
            write_block_file(d, lb0 + b, src + blk*b, (out / block_file(basename, lb0 + b)).string(), names);
    }
}

size_t image_doubles(const Grid& g)
{
    return 5*(size_t)g.geom.N[0]*g.geom.N[1]*g.geom.N[2]*g.geom.nblk;
}

} // namespace

namespace spd {
namespace {

// The context's two pinned checkpoint buffers, each at least n doubles.
void ensure_io_buffers(Ctx& c, size_t n)
{
    if (c.h_io_doubles < n)
    {
        for (int b = 0; b < 2; ++b)
        {
            if (c.h_io[b]) { SPD_HIP(hipHostFree(c.h_io[b])); c.h_io[b] = nullptr; }
            SPD_HIP(hipHostMalloc((void**)&c.h_io[b], sizeof(double)*n));
        }
        c.h_io_doubles = n;
    }
    for (int b = 0; b < 2; ++b)
        if (!c.io_ev[b]) SPD_HIP(hipEventCreateWithFlags(&c.io_ev[b], hipEventDisableTiming));
}

// Waits for the context's stream on every way out of a checkpoint call, so
// no copy into or out of a pinned buffer is still running when it returns.
struct StreamDrain
{
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

} // namespace
} // namespace spd

extern "C" {

int spd_io_output_vtk(spd_array_t arr, const char* out_dir, const char* basename, const char* const* names)
{
    return guarded([&]
    {
        require(arr != nullptr, "null argument");
        const Grid& g = *arr->a.grid;
        const size_t blk = 5*(size_t)g.geom.N[0]*g.geom.N[1]*g.geom.N[2];
        const int chunk = std::max(1, std::min(g.geom.nblk, (int)((64u << 20)/(8*blk))));
        output_vtk(g.desc, out_dir, basename, names, chunk, [&](int lb0, int nb, std::vector<double>& buf)
        {
            buf.resize(blk*nb);
            download_blocks(arr, lb0, nb, buf.data());
            return (const double*)buf.data();
        });
    });
}

int spd_io_output_vtk_host(const spd_grid_desc* desc, const double* host, const char* out_dir, const char* basename,
                           const char* const* names)
{
    return guarded([&]
    {
        validate_grid_desc(desc);
        require(host != nullptr, "null argument");
        const spd_grid_desc& d = *desc;
        const size_t blk = 5*(size_t)(d.cells_in_block[0] + 2*d.num_exch)*(d.cells_in_block[1] + 2*d.num_exch)*
                           (d.cells_in_block[2] + 2*d.num_exch);
        const int nblk = d.num_blocks[0]*d.num_blocks[1]*d.num_blocks[2];
        output_vtk(d, out_dir, basename, names, nblk, [&](int lb0, int, std::vector<double>&)
        {
            return host + blk*lb0;
        });
    });
}

int spd_io_binary_write(spd_array_t arr, const char* filename)
{
    return guarded([&]
    {
        require(arr && filename, "null argument");
        Grid& g = *arr->a.grid;
        Ctx& cx = *g.ctx;
        const size_t blk = 5*(size_t)g.geom.N[0]*g.geom.N[1]*g.geom.N[2];
        const int chunk = std::max(1, std::min(g.geom.nblk, (int)((64u << 20)/(8*blk))));
        const int nchunks = (g.geom.nblk + chunk - 1)/chunk;
        std::ofstream f(filename, std::ios::binary | std::ios::trunc);
        if (!f) throw IoError{std::string("cannot open ") + filename + " for writing"};
        ensure_io_buffers(cx, blk*chunk);
        StreamDrain drain{cx.stream};
        // chunk c goes through pinned buffer c % 2: the device converts and
        // copies chunk c+1 while the host writes chunk c
        auto enqueue = [&](int c)
        {
            const int lb0 = c*chunk, nb = std::min(chunk, g.geom.nblk - lb0);
            download_blocks(arr, lb0, nb, cx.h_io[c & 1], false);
            SPD_HIP(hipEventRecord(cx.io_ev[c & 1], cx.stream));
        };
        enqueue(0);
        for (int c = 0; c < nchunks; ++c)
        {
            if (c + 1 < nchunks) enqueue(c + 1);
            SPD_HIP(hipEventSynchronize(cx.io_ev[c & 1]));
            const int nb = std::min(chunk, g.geom.nblk - c*chunk);
            f.write(reinterpret_cast<const char*>(cx.h_io[c & 1]), (std::streamsize)(sizeof(double)*blk*nb));
            if (!f) throw IoError{std::string("cannot write ") + filename};
        }
        f.close();
        if (!f) throw IoError{std::string("cannot write ") + filename};
    });
}

int spd_io_binary_read(spd_array_t arr, const char* filename)
{
    return guarded([&]
    {
        require(arr && filename, "null argument");
        Grid& g = *arr->a.grid;
        const size_t blk = 5*(size_t)g.geom.N[0]*g.geom.N[1]*g.geom.N[2];
        std::ifstream f(filename, std::ios::binary | std::ios::ate);
        if (!f) throw IoError{std::string("cannot open ") + filename + " for reading"};
        const std::streamoff size = f.tellg();
        if (size != (std::streamoff)(sizeof(double)*image_doubles(g)))
            throw IoError{std::string(filename) + " holds " + std::to_string((long long)size) + " bytes, the array image " +
                          std::to_string(sizeof(double)*image_doubles(g))};
        f.seekg(0);
        const int chunk = std::max(1, std::min(g.geom.nblk, (int)((64u << 20)/(8*blk))));
        Ctx& cx = *g.ctx;
        ensure_io_buffers(cx, blk*chunk);
        StreamDrain drain{cx.stream};
        // chunk c goes through pinned buffer c % 2: the host reads chunk c
        // while the device copies and converts chunk c-1
        for (int lb0 = 0, c = 0; lb0 < g.geom.nblk; lb0 += chunk, ++c)
        {
            const int nb = std::min(chunk, g.geom.nblk - lb0);
            if (c >= 2) SPD_HIP(hipEventSynchronize(cx.io_ev[c & 1]));     // chunk c-2 has left this buffer
            f.read(reinterpret_cast<char*>(cx.h_io[c & 1]), (std::streamsize)(sizeof(double)*blk*nb));
            if (!f) throw IoError{std::string("cannot read ") + filename};
            upload_blocks(arr, lb0, nb, cx.h_io[c & 1], false);
            SPD_HIP(hipEventRecord(cx.io_ev[c & 1], cx.stream));
        }
    });
}

} // extern "C"
