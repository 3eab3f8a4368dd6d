// MAT v4 files on the host (BioFVM_matlab.h): the reader and writers of the
// reference's public API, which its example programs use for inputs and
// outputs beside Microenvironment::write_to_matlab.  Level-4 format: a
// 20-byte header (type = 1000*endianness + 100*0 + 10*data format + matrix
// type, rows, columns, imaginary flag, name length), the name, then the
// matrix in column-major order.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "BioFVM.h"

namespace BioFVM {

namespace {
// Opens `filename` and reads its header up to the data: the data format
// (0 double, 1 float, 2 int32, 3 int16, 4 uint16, 5 uint8) in *format; null
// (after the reference's message) when the file is missing or not a
// little-endian, real, full level-4 matrix (BioFVM_matlab.cpp:51-112).
FILE* open_level4(const std::string& filename, unsigned* rows, unsigned* cols, unsigned* format,
                  std::string* name)
{
    FILE* fp = std::fopen(filename.c_str(), "rb");
    if (!fp) {
        std::cout << "Error: could not open file " << filename << "!" << std::endl;
        return nullptr;
    }
    unsigned h[5] = {0, 0, 0, 0, 0};
    const bool ok = std::fread(h, sizeof(unsigned), 5, fp) == 5;
    const unsigned type = h[0];
    if (!ok || type / 1000 != 0 || (type / 100) % 10 != 0 || (type / 10) % 10 > 5 || type % 10 != 0) {
        std::cout << "Error reading file " << filename << ": I can't read this format yet!" << std::endl;
        std::fclose(fp);
        return nullptr;
    }
    if (h[3] != 0) {
        std::cout << "Error: I can't read imaginary matrices yet!" << std::endl;
        std::fclose(fp);
        return nullptr;
    }
    std::string nm(h[4], '\0');
    if (h[4] && std::fread(&nm[0], 1, h[4], fp) != h[4]) {
        std::fclose(fp);
        return nullptr;
    }
    *rows = h[1];
    *cols = h[2];
    *format = (type / 10) % 10;
    if (name) *name = nm.c_str();  // the stored name may end in a NUL
    return fp;
}

template <typename T>
bool read_column_major(FILE* fp, unsigned rows, unsigned cols, std::vector<std::vector<double> >& out)
{
    std::vector<T> buf((size_t)rows);
    for (unsigned j = 0; j < cols; ++j) {
        if (rows && std::fread(buf.data(), sizeof(T), rows, fp) != rows) return false;
        for (unsigned i = 0; i < rows; ++i) out[i][j] = (double)buf[i];
    }
    return true;
}
}  // namespace

std::vector<std::vector<double> > read_matlab(std::string filename)
{
    std::vector<std::vector<double> > out;
    unsigned rows = 0, cols = 0, format = 0;
    FILE* fp = open_level4(filename, &rows, &cols, &format, nullptr);
    if (!fp) return out;
    out.assign(rows, std::vector<double>(cols, 0.0));
    bool ok = false;
    switch (format) {
    case 0: ok = read_column_major<double>(fp, rows, cols, out); break;
    case 1: ok = read_column_major<float>(fp, rows, cols, out); break;
    case 2: ok = read_column_major<int>(fp, rows, cols, out); break;
    case 3: ok = read_column_major<short>(fp, rows, cols, out); break;
    case 4: ok = read_column_major<unsigned short>(fp, rows, cols, out); break;
    default: ok = read_column_major<unsigned char>(fp, rows, cols, out); break;
    }
    if (!ok) std::cout << "Error: " << filename << " ends before its matrix does" << std::endl;
    std::fclose(fp);
    return out;
}

FILE* read_matlab_header(int* rows, int* cols, std::string filename)
{
    unsigned r = 0, c = 0, format = 0;
    FILE* fp = open_level4(filename, &r, &c, &format, nullptr);
    if (!fp) return nullptr;
    *rows = (int)r;
    *cols = (int)c;
    return fp;
}

// BioFVM_matlab.cpp:515-564: a double, full, real matrix; the caller writes
// rows*cols doubles column by column and closes the file.
FILE* write_matlab_header(int rows, int cols, std::string filename, std::string variable_name)
{
    FILE* fp = std::fopen(filename.c_str(), "wb");
    if (!fp) {
        std::cout << "Error: could not open file " << filename << "!" << std::endl;
        return nullptr;
    }
    const unsigned h[5] = {0u, (unsigned)rows, (unsigned)cols, 0u, (unsigned)variable_name.size()};
    std::fwrite(h, sizeof(unsigned), 5, fp);
    std::fwrite(variable_name.data(), 1, variable_name.size(), fp);
    return fp;
}

// input[i] is column i (BioFVM_matlab.cpp:566-601), variable name "none".
bool write_matlab(std::vector<std::vector<double> >& input, std::string filename)
{
    const int cols = (int)input.size(), rows = cols ? (int)input[0].size() : 0;
    FILE* fp = write_matlab_header(rows, cols, filename, "none");
    if (!fp) return false;
    for (const std::vector<double>& col : input) std::fwrite(col.data(), sizeof(double), col.size(), fp);
    std::fclose(fp);
    return true;
}

}  // namespace BioFVM
