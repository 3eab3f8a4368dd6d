// Host helpers of BioFVM's public API that its example programs use beside
// the Microenvironment: element-wise std::vector<double> arithmetic and norms
// (BioFVM_vector.h/.cpp), and the stopwatch (BioFVM_utilities.h/.cpp).  Same
// names, signatures and results; none of it touches the device.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <string>
#include <vector>

#include "BioFVM.h"

namespace BioFVM {

namespace {
template <class F>
std::vector<double> zip(const std::vector<double>& a, const std::vector<double>& b, F f)
{
    std::vector<double> out(a);
    for (size_t i = 0; i < a.size(); ++i) out[i] = f(a[i], b[i]);
    return out;
}
template <class F>
std::vector<double> each(const std::vector<double>& a, F f)
{
    std::vector<double> out(a);
    for (double& x : out) x = f(x);
    return out;
}
}  // namespace

// Element-wise (Hadamard) arithmetic, BioFVM_vector.cpp:37-150.
std::vector<double> operator-(const std::vector<double>& v1, const std::vector<double>& v2)
{
    return zip(v1, v2, [](double a, double b) { return a - b; });
}
std::vector<double> operator+(const std::vector<double>& v1, const std::vector<double>& v2)
{
    return zip(v1, v2, [](double a, double b) { return a + b; });
}
std::vector<double> operator*(const std::vector<double>& v1, const std::vector<double>& v2)
{
    return zip(v1, v2, [](double a, double b) { return a * b; });
}
std::vector<double> operator/(const std::vector<double>& v1, const std::vector<double>& v2)
{
    return zip(v1, v2, [](double a, double b) { return a / b; });
}
std::vector<double> operator*(double d, const std::vector<double>& v1)
{
    return each(v1, [d](double x) { return d * x; });
}
std::vector<double> operator+(double d, const std::vector<double>& v1)
{
    return each(v1, [d](double x) { return d + x; });
}
std::vector<double> operator+(const std::vector<double>& v1, double d)
{
    return each(v1, [d](double x) { return x + d; });
}
std::vector<double> operator-(double d, const std::vector<double>& v1)
{
    return each(v1, [d](double x) { return d - x; });
}
std::vector<double> operator-(const std::vector<double>& v1, double d)
{
    return each(v1, [d](double x) { return x - d; });
}
void operator+=(std::vector<double>& v1, const std::vector<double>& v2)
{
    for (size_t i = 0; i < v1.size(); ++i) v1[i] += v2[i];
}
void operator-=(std::vector<double>& v1, const std::vector<double>& v2)
{
    for (size_t i = 0; i < v1.size(); ++i) v1[i] -= v2[i];
}
void operator/=(std::vector<double>& v1, const std::vector<double>& v2)
{
    for (size_t i = 0; i < v1.size(); ++i) v1[i] /= v2[i];
}
void operator*=(std::vector<double>& v1, const double& a)
{
    for (double& x : v1) x *= a;
}
void operator*=(std::vector<double>& v1, const std::vector<double>& v2)
{
    for (size_t i = 0; i < v1.size(); ++i) v1[i] *= v2[i];
}
void operator/=(std::vector<double>& v1, const double& a)
{
    for (double& x : v1) x /= a;
}

// Three values print as XML-style coordinates, others space-separated
// (BioFVM_vector.cpp:153-164).
std::ostream& operator<<(std::ostream& os, const std::vector<double>& v)
{
    if (v.size() == 3) return os << "x=\"" << v[0] << "\" y=\"" << v[1] << "\" z=\"" << v[2] << "\"";
    for (double x : v) os << x << " ";
    return os;
}

double norm_squared(const std::vector<double>& v)
{
    double out = 0.0;
    for (double x : v) out += x * x;
    return out;
}
double norm(const std::vector<double>& v) { return std::sqrt(norm_squared(v)); }

// A normalised copy (no guard: a zero vector gives NaNs, as the reference's).
std::vector<double> normalize(std::vector<double>& v)
{
    const double n = norm(v);
    return each(v, [n](double x) { return x / n; });
}
// In place, with 1e-32 added to the squared norm (BioFVM_vector.cpp:184-196).
void normalize(std::vector<double>* v)
{
    double n2 = 1e-32;
    for (double x : *v) n2 += x * x;
    const double n = std::sqrt(n2);
    for (double& x : *v) x /= n;
}

// The reference keeps the signed entry as its running value
// (BioFVM_vector.cpp:210-219): the result is the last entry whose magnitude
// exceeded the value kept before it, with its sign.
double maxabs(const std::vector<double>& v)
{
    double out = 0.0;
    for (double x : v)
        if (std::fabs(x) > out) out = x;
    return out;
}
double max_abs_difference(const std::vector<double>& v1, const std::vector<double>& v2)
{
    double out = 0.0;
    for (size_t i = 0; i < v1.size(); ++i) out = std::fmax(out, std::fabs(v1[i] - v2[i]));
    return out;
}
std::vector<double> exponentiate(const std::vector<double>& exponent)
{
    return each(exponent, [](double x) { return std::exp(x); });
}
// Uniform in [-1, 1] from rand() (BioFVM_vector.cpp:242-252).
void randomize(std::vector<double>* v)
{
    const double d1 = 2.0 / (double)RAND_MAX;
    for (double& x : *v) x = -1 + d1 * rand();
}

// y += a x, y -= a x, scalar or element-wise a (BioFVM_vector.cpp:256-290).
void axpy(std::vector<double>* y, double& a, std::vector<double>& x)
{
    for (size_t i = 0; i < y->size(); ++i) (*y)[i] += a * x[i];
}
void axpy(std::vector<double>* y, std::vector<double>& a, std::vector<double>& x)
{
    for (size_t i = 0; i < y->size(); ++i) (*y)[i] += a[i] * x[i];
}
void naxpy(std::vector<double>* y, double& a, std::vector<double>& x)
{
    for (size_t i = 0; i < y->size(); ++i) (*y)[i] -= a * x[i];
}
void naxpy(std::vector<double>* y, std::vector<double>& a, std::vector<double>& x)
{
    for (size_t i = 0; i < y->size(); ++i) (*y)[i] -= a[i] * x[i];
}

// Numbers of a delimited character array: skip to the next character that can
// start a number (digit, '.', '-', 'e', 'E'), read it with strtod
// (BioFVM_vector.cpp:294-311).
void csv_to_vector(const char* buffer, std::vector<double>& vect)
{
    vect.clear();
    const size_t n = std::strlen(buffer);
    size_t i = 0;
    while (i < n) {
        while (i < n && !std::isdigit((unsigned char)buffer[i]) && buffer[i] != '.' && buffer[i] != '-' &&
               buffer[i] != 'e' && buffer[i] != 'E')
            ++i;
        if (i >= n) break;
        char* end = nullptr;
        vect.push_back(std::strtod(buffer + i, &end));
        const size_t next = (size_t)(end - buffer);
        i = next > i ? next : i + 1;
    }
}

// "%.7e" values separated by commas; the caller owns the new[] buffer, 16
// characters per value as the reference allocates (BioFVM_vector.cpp:313-364).
void vector_to_csv(const std::vector<double>& vect, char*& buffer)
{
    int pos = 0;
    for (size_t j = 0; j < vect.size(); ++j)
        pos += std::sprintf(buffer + pos, j + 1 < vect.size() ? "%.7e," : "%.7e", vect[j]);
}
char* vector_to_csv(const std::vector<double>& vect)
{
    char* buffer = new char[16 * vect.size() + 16];
    vector_to_csv(vect, buffer);
    return buffer;
}
void vector_to_csv_safe(const std::vector<double>& vect, char*& buffer)
{
    delete[] buffer;
    buffer = new char[16 * vect.size() + 16];
    vector_to_csv(vect, buffer);
}

// Stopwatch (BioFVM_utilities.cpp:47-110): processor time from clock(), as the
// reference measures it, summed over TIC/TOC pairs.
namespace {
clock_t g_tic = 0, g_toc = 0, g_run_start = 0, g_run_end = 0;
double g_total = 0.0;
}  // namespace

void TIC(void) { g_tic = clock(); }
void TOC(void)
{
    g_toc = clock();
    g_total += stopwatch_value();
}
void RUNTIME_TIC(void) { g_run_start = clock(); }
void RUNTIME_TOC(void) { g_run_end = clock(); }
double stopwatch_value(void) { return (double)(g_toc - g_tic) / (double)CLOCKS_PER_SEC; }
double runtime_stopwatch_value(void) { return (double)(g_run_end - g_run_start) / (double)CLOCKS_PER_SEC; }
double total_stopwatch_time(void) { return g_total; }

namespace {
void split_seconds(double t, int& days, int& hours, int& minutes, double& seconds)
{
    days = (int)std::floor(t / 86400.0);
    hours = (int)std::floor((t - days * 86400.0) / 3600.0);
    minutes = (int)std::floor((t - days * 86400.0 - hours * 3600.0) / 60.0);
    seconds = t - days * 86400.0 - hours * 3600.0 - minutes * 60.0;
}
}  // namespace

void display_stopwatch_value(std::ostream& os, double dIn)
{
    int d, h, m;
    double s;
    split_seconds(dIn, d, h, m, s);
    os << d << " days, " << h << " hours, " << m << " minutes, and " << s << " seconds ";
}

std::string format_stopwatch_value(double dIn)
{
    int d, h, m;
    double s;
    split_seconds(dIn, d, h, m, s);
    char buf[128];
    std::snprintf(buf, sizeof buf, "%d days, %d hours, %d minutes, and %2.4f seconds", d, h, m, s);
    return buf;
}

}  // namespace BioFVM
