#include "Output.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>

#include "Run.h"

namespace dsmc_host {

namespace {

uint32_t crc32(const std::string &s)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (unsigned char b : s) c = table[(c ^ b) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

void put16(std::string &o, uint32_t v)
{
    o.push_back((char)(v & 0xFF));
    o.push_back((char)((v >> 8) & 0xFF));
}

void put32(std::string &o, uint32_t v)
{
    put16(o, v & 0xFFFF);
    put16(o, v >> 16);
}

// One .npy member: format 1.0, little-endian, C order, header padded so the
// data starts on a 64-byte boundary.
template <class T>
std::string npy(const std::vector<T> &a, int n1, int n2, const char *descr)
{
    std::string h = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (" +
                    std::to_string(n1) + ", " + std::to_string(n2) + "), }";
    size_t total = 10 + h.size() + 1;
    h.append((64 - total % 64) % 64, ' ');
    h.push_back('\n');
    std::string o("\x93NUMPY\x01\x00", 8);
    put16(o, (uint32_t)h.size());
    o += h;
    o.append(reinterpret_cast<const char *>(a.data()), sizeof(T) * a.size());
    return o;
}

std::string fmt9(double v)
{
    if (std::isnan(v)) return "nan";
    char b[40];
    std::snprintf(b, sizeof b, "%.9g", v);
    return b;
}

}  // namespace

void Output::write_npz(const std::string &path, const FieldSet &f)
{
    struct Member { std::string name, data; };
    std::vector<Member> m;
    const char *names[] = { "U", "V", "W", "P", "T", "Rho" };
    const std::vector<double> *vals[] = { &f.U, &f.V, &f.W, &f.P, &f.T, &f.Rho };
    for (int k = 0; k < 6; ++k) m.push_back({ std::string(names[k]) + ".npy", npy(*vals[k], f.N1, f.N2, "<f8") });
    m.push_back({ "Numchild.npy", npy(f.Numchild, f.N1, f.N2, "<i4") });
    m.push_back({ "Numgrandchild.npy", npy(f.Numgrandchild, f.N1, f.N2, "<i4") });
    m.push_back({ "MFP.npy", npy(f.MFP, f.N1, f.N2, "<f8") });
    if (!f.Trot.empty()) m.push_back({ "Trot.npy", npy(f.Trot, f.N1, f.N2, "<f8") });

    // zip archive, every member stored (no compression)
    std::string out, cd;
    for (const Member &e : m) {
        const uint32_t crc = crc32(e.data), size = (uint32_t)e.data.size(), off = (uint32_t)out.size();
        put32(out, 0x04034b50u);
        put16(out, 20); put16(out, 0); put16(out, 0); put16(out, 0); put16(out, 0x21);
        put32(out, crc); put32(out, size); put32(out, size);
        put16(out, (uint32_t)e.name.size()); put16(out, 0);
        out += e.name;
        out += e.data;
        put32(cd, 0x02014b50u);
        put16(cd, 20); put16(cd, 20); put16(cd, 0); put16(cd, 0); put16(cd, 0); put16(cd, 0x21);
        put32(cd, crc); put32(cd, size); put32(cd, size);
        put16(cd, (uint32_t)e.name.size()); put16(cd, 0); put16(cd, 0); put16(cd, 0); put16(cd, 0);
        put32(cd, 0); put32(cd, off);
        cd += e.name;
    }
    const uint32_t cd_off = (uint32_t)out.size();
    out += cd;
    put32(out, 0x06054b50u);
    put16(out, 0); put16(out, 0);
    put16(out, (uint32_t)m.size()); put16(out, (uint32_t)m.size());
    put32(out, (uint32_t)cd.size()); put32(out, cd_off);
    put16(out, 0);
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error("cannot write " + path);
    const size_t w = std::fwrite(out.data(), 1, out.size(), fp);
    std::fclose(fp);
    if (w != out.size()) throw std::runtime_error("short write to " + path);
}

void Output::write_vts(const std::string &path, const FieldSet &f)
{
    const int n1 = f.N1, n2 = f.N2;
    const double dx = f.L1 / n1, dy = f.L2 / n2;
    std::string s;
    s += "<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    s += "<StructuredGrid WholeExtent=\"0 " + std::to_string(n1) + " 0 " + std::to_string(n2) + " 0 0\">\n";
    s += "<Piece Extent=\"0 " + std::to_string(n1) + " 0 " + std::to_string(n2) + " 0 0\">\n";
    s += "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int j = 0; j <= n2; ++j) {
        for (int i = 0; i <= n1; ++i) {
            if (i) s += ' ';
            s += fmt9(i * dx) + ' ' + fmt9(j * dy) + " 0";
        }
        s += '\n';
    }
    s += "</DataArray>\n</Points>\n<CellData>\n";
    auto array = [&](const char *name, auto get, bool is_int) {
        s += std::string("<DataArray type=\"") + (is_int ? "Int32" : "Float64") + "\" Name=\"" + name +
             "\" format=\"ascii\">\n";
        const size_t n = (size_t)n1 * n2;
        for (size_t q = 0; q < n; ++q) {
            const int j = (int)(q / n1), i = (int)(q % n1);     // x fastest
            const size_t c = (size_t)i * n2 + j;
            if (q % 512) s += ' ';
            s += get(c);
            if (q % 512 == 511 || q + 1 == n) s += '\n';
        }
        s += "</DataArray>\n";
    };
    const char *names[] = { "U", "V", "W", "P", "T", "Rho" };
    const std::vector<double> *vals[] = { &f.U, &f.V, &f.W, &f.P, &f.T, &f.Rho };
    for (int k = 0; k < 6; ++k) array(names[k], [&](size_t c) { return fmt9((*vals[k])[c]); }, false);
    array("Numchild", [&](size_t c) { return std::to_string(f.Numchild[c]); }, true);
    array("Numgrandchild", [&](size_t c) { return std::to_string(f.Numgrandchild[c]); }, true);
    array("MFP", [&](size_t c) { return fmt9(f.MFP[c]); }, false);
    if (!f.Trot.empty()) array("Trot", [&](size_t c) { return fmt9(f.Trot[c]); }, false);
    s += "</CellData>\n</Piece>\n</StructuredGrid>\n</VTKFile>\n";
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) throw std::runtime_error("cannot write " + path);
    const size_t w = std::fwrite(s.data(), 1, s.size(), fp);
    std::fclose(fp);
    if (w != s.size()) throw std::runtime_error("short write to " + path);
}

void Output::Write2HDF5(const std::string &filename)
{
    write_npz(filename + ".npz", m_run->fields());
}

void Output::Write2VTK(const std::string &filename)
{
    write_vts(filename + ".vts", m_run->fields());
}

}  // namespace dsmc_host
