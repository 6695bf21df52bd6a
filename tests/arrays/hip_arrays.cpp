// The reference's compressed arrays (zfp::array1..4 of LLNL zfp 1.0.1,
// include/zfp/array*.hpp) with the GPU block codecs of
// include/zfp/hipcodec.hpp, against the same arrays with their default
// codecs: identical compressed data (byte for byte) and identical values
// from get(), element reads, and after element writes and cache flushes,
// with and without a batch scope.  Built by oracle/Makefile against the
// reference's headers and library; run by tests/test_gpu_arrays.py.
// Prints one line per case ("ok ..." or "FAIL ...") and exits non-zero on
// any failure; with --time, also times set() and get() of a 256^3 array.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "zfp/array1.hpp"
#include "zfp/array2.hpp"
#include "zfp/array3.hpp"
#include "zfp/array4.hpp"
#include "zfp/hipcodec.hpp"

static int failures = 0;

static void report(bool ok, const std::string& what)
{
  std::printf("%s %s\n", ok ? "ok" : "FAIL", what.c_str());
  std::fflush(stdout);
  if (!ok)
    failures++;
}

template <typename T>
static std::vector<T> field(size_t n, unsigned seed)
{
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(-1, 1);
  std::vector<T> v(n);
  double a = 0;
  for (size_t i = 0; i < n; i++) {
    a = 0.97 * a + 0.1 * u(g);
    v[i] = (T)(std::sin(0.01 * (double)i) + a);
  }
  return v;
}

template <class A, class B>
static bool same_data(const A& a, const B& b)
{
  return a.compressed_size() == b.compressed_size() &&
         !std::memcmp(a.compressed_data(), b.compressed_data(), a.compressed_size());
}

template <typename T>
static bool same_values(const std::vector<T>& x, const std::vector<T>& y)
{
  return x.size() == y.size() && !std::memcmp(x.data(), y.data(), x.size() * sizeof(T));
}

// the array-level checks for one shape, rate and scalar type; Ref and Hip
// are array classes of the same dimensionality with the default and the GPU
// codec, Make builds one from (rate, data)
template <typename T, class Ref, class Hip, class MakeRef, class MakeHip, class Touch>
static void check(const std::string& name, size_t n, double rate, MakeRef make_ref, MakeHip make_hip, Touch touch)
{
  const std::vector<T> v = field<T>(n, (unsigned)(n * 31 + (size_t)rate));
  Ref ref = make_ref(rate, v.data());
  Hip one = make_hip(rate, v.data());          // block by block
  Hip many = make_hip(rate, nullptr);
  {
    zfp::codec::hip_batch batch;               // all blocks in one launch
    many.set(v.data());
  }
  report(same_data(ref, one), name + ": set() block by block, same compressed data");
  report(same_data(ref, many), name + ": set() in a batch, same compressed data");
  std::vector<T> gr(n), g1(n), gm(n);
  ref.get(gr.data());
  one.get(g1.data());
  {
    zfp::codec::hip_batch batch;
    many.get(gm.data());
  }
  report(same_values(gr, g1), name + ": get() block by block");
  report(same_values(gr, gm), name + ": get() in a batch");
  // element reads (cache fills), then writes and a flush (cache write-back)
  bool reads = true;
  for (size_t i = 0; i < n; i += 1 + n / 97) {
    const T a = ref[i], b = one[i], c = many[i];
    reads = reads && !std::memcmp(&a, &b, sizeof(T)) && !std::memcmp(&a, &c, sizeof(T));
  }
  report(reads, name + ": element reads");
  touch(ref);
  touch(one);
  {
    zfp::codec::hip_batch batch;
    touch(many);
    many.flush_cache();                        // write-back queued
  }
  ref.flush_cache();
  one.flush_cache();
  report(same_data(ref, one), name + ": element writes + flush, block by block");
  report(same_data(ref, many), name + ": element writes + flush in a batch");
  ref.get(gr.data());
  {
    zfp::codec::hip_batch batch;
    many.get(gm.data());
  }
  report(same_values(gr, gm), name + ": get() after writes");
}

template <typename T>
static void run_type(const char* tname)
{
  typedef zfp::array1<T, zfp::codec::zfp1<T>> R1;
  typedef zfp::array1<T, zfp::codec::hip1<T>> H1;
  typedef zfp::array2<T, zfp::codec::zfp2<T>> R2;
  typedef zfp::array2<T, zfp::codec::hip2<T>> H2;
  typedef zfp::array3<T, zfp::codec::zfp3<T>> R3;
  typedef zfp::array3<T, zfp::codec::hip3<T>> H3;
  typedef zfp::array4<T, zfp::codec::zfp4<T>> R4;
  typedef zfp::array4<T, zfp::codec::hip4<T>> H4;
  const double rates[] = {5, 8, 16, 32};
  for (double rate : rates) {
    char tag[64];
    const size_t s3[][3] = {{16, 16, 16}, {13, 10, 7}, {65, 33, 18}};
    for (auto& s : s3) {
      std::snprintf(tag, sizeof tag, "3D %s %zux%zux%zu rate %g", tname, s[0], s[1], s[2], rate);
      check<T, R3, H3>(tag, s[0] * s[1] * s[2], rate,
                       [&](double r, const T* p) { return R3(s[0], s[1], s[2], r, p); },
                       [&](double r, const T* p) { return H3(s[0], s[1], s[2], r, p); },
                       [&](auto& a) {
                         a(0, 0, 0) = (T)1.5;
                         a(s[0] - 1, s[1] - 1, s[2] - 1) = (T)-2.25;
                         a(s[0] / 2, s[1] / 3, s[2] / 2) += (T)0.125;
                       });
    }
    std::snprintf(tag, sizeof tag, "2D %s 37x29 rate %g", tname, rate);
    check<T, R2, H2>(tag, 37 * 29, rate, [&](double r, const T* p) { return R2(37, 29, r, p); },
                     [&](double r, const T* p) { return H2(37, 29, r, p); },
                     [&](auto& a) { a(3, 4) = (T)7; a(36, 28) = (T)-1; });
    std::snprintf(tag, sizeof tag, "1D %s 1001 rate %g", tname, rate);
    check<T, R1, H1>(tag, 1001, rate, [&](double r, const T* p) { return R1(1001, r, p); },
                     [&](double r, const T* p) { return H1(1001, r, p); },
                     [&](auto& a) { a(0) = (T)3; a(1000) = (T)-3; });
    std::snprintf(tag, sizeof tag, "4D %s 9x6x5x7 rate %g", tname, rate);
    check<T, R4, H4>(tag, 9 * 6 * 5 * 7, rate, [&](double r, const T* p) { return R4(9, 6, 5, 7, r, p); },
                     [&](double r, const T* p) { return H4(9, 6, 5, 7, r, p); },
                     [&](auto& a) { a(1, 2, 3, 4) = (T)0.5; a(8, 5, 4, 6) = (T)9; });
  }
}

static void time_256()
{
  typedef zfp::array3<double, zfp::codec::zfp3<double>> R3;
  typedef zfp::array3<double, zfp::codec::hip3<double>> H3;
  const size_t n = 256;
  const std::vector<double> v = field<double>(n * n * n, 7);
  std::vector<double> out(n * n * n);
  for (double rate : {8.0, 16.0}) {
    R3 ref(n, n, n, rate);
    H3 hip(n, n, n, rate);
    auto t = [] { return std::chrono::steady_clock::now(); };
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    {  // first GPU call of the process
      zfp::codec::hip_batch batch;
      hip.set(v.data());
    }
    auto t0 = t();
    ref.set(v.data());
    auto t1 = t();
    {
      zfp::codec::hip_batch batch;
      hip.set(v.data());
    }
    auto t2 = t();
    ref.get(out.data());
    auto t3 = t();
    {
      zfp::codec::hip_batch batch;
      hip.get(out.data());
    }
    auto t4 = t();
    std::printf("time 3D double 256^3 rate %g: set() %.1f ms serial, %.1f ms GPU batch; get() %.1f ms serial, "
                "%.1f ms GPU batch; same data %d\n", rate, ms(t0, t1), ms(t1, t2), ms(t2, t3), ms(t3, t4),
                (int)same_data(ref, hip));
  }
}

int main(int argc, char** argv)
{
  try {
    run_type<double>("double");
    run_type<float>("float");
    if (argc > 1 && !std::strcmp(argv[1], "--time"))
      time_256();
  }
  catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 2;
  }
  std::printf("%d failures\n", failures);
  return failures ? 1 : 0;
}
