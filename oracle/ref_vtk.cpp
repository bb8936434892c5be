// Test infrastructure only: drives VirtualFluids' own binary VTK XML writer
// (src/basics/writer/WbWriterVtkXmlBinary.cpp, compiled from the reference
// sources where they lie) to write the files the output tests pin byte for
// byte.  Modes:
//   ref_vtk octs <input> <prefix>   writeOctsWithNodeData (:1306-1425); input:
//       int32 n, float32 nodes[3n], int32 m, uint32 cells[8m], int32 k, then k
//       times { int32 length, name bytes, float64 values[n] }
//   ref_vtk pvtu <prefix> <names,comma,separated> <piece>...   writeParallelFile
//   ref_vtk pvd <prefix> <step>=<file>[,<file>...] ...         writeCollectionForTimeSeries
//   ref_vtk vti <input> <file>   WbWriterVtkXmlImageBinary::writeData (the
//       precursor planes of PrecursorWriter::write); input: int32 k, k times
//       { int32 length, name bytes }, int32 n, float64 values[k][n],
//       int32 extent[6], float32 origin[3], float32 spacing[3], int32 precision
// Prints the path the writer returns (vti: the file name).
#include <basics/writer/WbWriterVtkXmlBinary.h>
#include <basics/writer/WbWriterVtkXmlImageBinary.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

static std::vector<std::string> split(const std::string& s, char sep)
{
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, sep))
        if (!item.empty()) out.push_back(item);
    return out;
}

template <typename T> static T get(std::ifstream& in)
{
    T v{};
    in.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!in) throw std::runtime_error("short input");
    return v;
}

static int octs(const char* input, const std::string& prefix)
{
    std::ifstream in(input, std::ios::binary);
    if (!in) return 2;
    const int32_t n = get<int32_t>(in);
    std::vector<UbTupleFloat3> nodes;
    nodes.reserve(n);
    for (int32_t i = 0; i < n; ++i) {
        const float x = get<float>(in), y = get<float>(in), z = get<float>(in);
        nodes.push_back(makeUbTuple(x, y, z));
    }
    const int32_t m = get<int32_t>(in);
    std::vector<UbTupleUInt8> cells;
    cells.reserve(m);
    for (int32_t c = 0; c < m; ++c) {
        unsigned int v[8];
        for (auto& x : v) x = get<uint32_t>(in);
        cells.push_back(makeUbTuple(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]));
    }
    const int32_t k = get<int32_t>(in);
    std::vector<std::string> names;
    std::vector<std::vector<double>> data;
    for (int32_t j = 0; j < k; ++j) {
        const int32_t len = get<int32_t>(in);
        std::string name(len, '\0');
        in.read(&name[0], len);
        std::vector<double> d(n);
        for (auto& x : d) x = get<double>(in);
        names.push_back(name);
        data.push_back(std::move(d));
    }
    std::printf("%s\n", WbWriterVtkXmlBinary::getInstance()->writeOctsWithNodeData(prefix, nodes, cells, names, data).c_str());
    return 0;
}

static int vti(const char* input, const std::string& file)
{
    std::ifstream in(input, std::ios::binary);
    if (!in) return 2;
    const int32_t k = get<int32_t>(in);
    std::vector<std::string> names;
    for (int32_t j = 0; j < k; ++j) {
        const int32_t len = get<int32_t>(in);
        std::string name(len, '\0');
        in.read(&name[0], len);
        names.push_back(name);
    }
    const int32_t n = get<int32_t>(in);
    std::vector<std::vector<double>> data(k, std::vector<double>(n));
    for (auto& d : data)
        for (auto& x : d) x = get<double>(in);
    int32_t e[6];
    for (auto& x : e) x = get<int32_t>(in);
    float o[3], sp[3];
    for (auto& x : o) x = get<float>(in);
    for (auto& x : sp) x = get<float>(in);
    const int32_t precision = get<int32_t>(in);
    UbTupleInt6 extent = makeUbTuple(e[0], e[1], e[2], e[3], e[4], e[5]);
    UbTupleFloat3 origin = makeUbTuple(o[0], o[1], o[2]);
    UbTupleFloat3 spacing = makeUbTuple(sp[0], sp[1], sp[2]);
    std::vector<std::string> cellNames;
    std::vector<std::vector<double>> cellData;
    WbWriterVtkXmlImageBinary::getInstance()->writeData(file, names, cellNames, data, cellData, extent, origin, spacing,
                                                        extent, (unsigned)precision);
    std::printf("%s\n", file.c_str());
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "octs" && argc == 4) return octs(argv[2], argv[3]);
        if (mode == "vti" && argc == 4) return vti(argv[2], argv[3]);
        if (mode == "pvtu" && argc >= 4) {
            std::vector<std::string> names = split(argv[3], ','), cells, pieces;
            for (int i = 4; i < argc; ++i) pieces.push_back(argv[i]);
            std::printf("%s\n", WbWriterVtkXmlBinary::getInstance()->writeParallelFile(argv[2], pieces, names, cells).c_str());
            return 0;
        }
        if (mode == "pvd") {
            std::map<uint, std::vector<std::string>> files;
            for (int i = 3; i < argc; ++i) {
                const std::string a = argv[i];
                const size_t eq = a.find('=');
                if (eq == std::string::npos) return 2;
                files[(uint)std::strtoul(a.substr(0, eq).c_str(), nullptr, 10)] = split(a.substr(eq + 1), ',');
            }
            std::printf("%s\n",
                        WbWriterVtkXmlBinary::getInstance()->writeCollectionForTimeSeries(argv[2], files, false).c_str());
            return 0;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 2;
}
