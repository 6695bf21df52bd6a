/*
 * zfp_cli.c -- the `zfp` command-line tool over libzfp_amd, with the
 * execution policy zfp_exec_hip selectable by `-x hip`.
 *
 * Options, file roles, messages and exit codes follow the reference tool
 * (utils/zfp.c:83-135 usage, :176-292 options, :294-353 checks, :355-629
 * tasks):
 *   -i <path>  uncompressed input ("-" for stdin): read and compress
 *   -z <path>  compressed output with -i, compressed input without it
 *   -o <path>  decompressed output ("-" for stdout)
 *   -h         full zfp header before the stream (zfp_write_header /
 *              zfp_read_header, mask ZFP_HEADER_FULL)
 *   -s         error statistics (needs -i); -q quiet
 *   -f -d -t <i32|i64|f32|f64>; -1 .. -4 <extents>; -R -r -p -a -c
 *   -x serial | omp[=threads[,chunk]] | cuda | hip
 * One difference: the default policy is hip, the one this library computes
 * with.  serial, omp and cuda are accepted as in the reference; this library
 * has no CPU or CUDA codec, so compression under them fails with the
 * reference's "compression failed".
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zfp.h"

typedef struct {
  zfp_type type;
  uint dims;
  size_t n[4];
  char mode; /* 'R', 'r', 'p', 'a', 'c' or 0 */
  double rate, tolerance;
  uint precision;
  uint minbits, maxbits, maxprec;
  int minexp;
  zfp_exec_policy exec;
  uint threads, chunk;
  int header, quiet, stats;
  const char *in, *zfp, *out;
} options;

static void usage(void)
{
  static const char* text[] = {
    "Usage: zfp <options>",
    "General options:",
    "  -h : read/write array and compression parameters from/to compressed header",
    "  -q : quiet mode; suppress output",
    "  -s : print error statistics",
    "Input and output:",
    "  -i <path> : uncompressed binary input file (\"-\" for stdin)",
    "  -o <path> : decompressed binary output file (\"-\" for stdout)",
    "  -z <path> : compressed input (w/o -i) or output file (\"-\" for stdin/stdout)",
    "Array type and dimensions (needed with -i):",
    "  -f : single precision (float type)",
    "  -d : double precision (double type)",
    "  -t <i32|i64|f32|f64> : integer or floating scalar type",
    "  -1 <nx> : dimensions for 1D array a[nx]",
    "  -2 <nx> <ny> : dimensions for 2D array a[ny][nx]",
    "  -3 <nx> <ny> <nz> : dimensions for 3D array a[nz][ny][nx]",
    "  -4 <nx> <ny> <nz> <nw> : dimensions for 4D array a[nw][nz][ny][nx]",
    "Compression parameters (needed with -i):",
    "  -R : reversible (lossless) compression",
    "  -r <rate> : fixed rate (# compressed bits per floating-point value)",
    "  -p <precision> : fixed precision (# uncompressed bits per value)",
    "  -a <tolerance> : fixed accuracy (absolute error tolerance)",
    "  -c <minbits> <maxbits> <maxprec> <minexp> : advanced usage",
    "Execution parameters:",
    "  -x hip : MI355X (gfx950) compression/decompression, all modes (default)",
    "  -x serial | omp[=threads[,chunk_size]] | cuda : the reference's policies,",
    "     accepted but not computed by this library",
    NULL};
  fprintf(stderr, "%s\n", zfp_version_string);
  for (int i = 0; text[i]; i++)
    fprintf(stderr, "%s\n", text[i]);
  exit(EXIT_FAILURE);
}

/* next argument as an unsigned size, a uint, an int or a double */
static const char* next_arg(int argc, char** argv, int* i)
{
  if (++*i >= argc)
    usage();
  return argv[*i];
}
static size_t arg_size(int argc, char** argv, int* i)
{
  size_t v;
  if (sscanf(next_arg(argc, argv, i), "%zu", &v) != 1)
    usage();
  return v;
}
static uint arg_uint(int argc, char** argv, int* i)
{
  uint v;
  if (sscanf(next_arg(argc, argv, i), "%u", &v) != 1)
    usage();
  return v;
}
static double arg_double(int argc, char** argv, int* i)
{
  double v;
  if (sscanf(next_arg(argc, argv, i), "%lf", &v) != 1)
    usage();
  return v;
}

static void parse(int argc, char** argv, options* o)
{
  memset(o, 0, sizeof(*o));
  o->type = zfp_type_none;
  o->minbits = ZFP_MIN_BITS;
  o->maxbits = ZFP_MAX_BITS;
  o->maxprec = ZFP_MAX_PREC;
  o->minexp = ZFP_MIN_EXP;
  o->exec = zfp_exec_hip;
  if (argc == 1)
    usage();
  for (int i = 1; i < argc; i++) {
    const char* a = argv[i];
    if (a[0] != '-' || !a[1] || a[2])
      usage();
    switch (a[1]) {
      case '1': case '2': case '3': case '4':
        o->dims = (uint)(a[1] - '0');
        for (uint d = 0; d < 4; d++)
          o->n[d] = d < o->dims ? arg_size(argc, argv, &i) : 1;
        break;
      case 'a': o->tolerance = arg_double(argc, argv, &i); o->mode = 'a'; break;
      case 'c':
        o->minbits = arg_uint(argc, argv, &i);
        o->maxbits = arg_uint(argc, argv, &i);
        o->maxprec = arg_uint(argc, argv, &i);
        if (sscanf(next_arg(argc, argv, &i), "%d", &o->minexp) != 1)
          usage();
        o->mode = 'c';
        break;
      case 'd': o->type = zfp_type_double; break;
      case 'f': o->type = zfp_type_float; break;
      case 'h': o->header = 1; break;
      case 'i': o->in = next_arg(argc, argv, &i); break;
      case 'o': o->out = next_arg(argc, argv, &i); break;
      case 'p': o->precision = arg_uint(argc, argv, &i); o->mode = 'p'; break;
      case 'q': o->quiet = 1; break;
      case 'r': o->rate = arg_double(argc, argv, &i); o->mode = 'r'; break;
      case 'R': o->mode = 'R'; break;
      case 's': o->stats = 1; break;
      case 't': {
        const char* t = next_arg(argc, argv, &i);
        static const char* names[] = {"i32", "i64", "f32", "f64"};
        o->type = zfp_type_none;
        for (int k = 0; k < 4; k++)
          if (!strcmp(t, names[k]))
            o->type = (zfp_type)(zfp_type_int32 + k);
        if (o->type == zfp_type_none)
          usage();
        break;
      }
      case 'x': {
        const char* x = next_arg(argc, argv, &i);
        uint t = 0, c = 0;
        if (!strcmp(x, "hip"))
          o->exec = zfp_exec_hip;
        else if (!strcmp(x, "serial"))
          o->exec = zfp_exec_serial;
        else if (!strcmp(x, "cuda"))
          o->exec = zfp_exec_cuda;
        else if (!strcmp(x, "omp") || sscanf(x, "omp=%u,%u", &t, &c) == 2 || sscanf(x, "omp=%u", &t) == 1) {
          o->exec = zfp_exec_omp;
          o->threads = t;
          o->chunk = c;
        }
        else
          usage();
        break;
      }
      case 'z': o->zfp = next_arg(argc, argv, &i); break;
      default: usage();
    }
  }
}

/* the reference's consistency checks (utils/zfp.c:294-353); 0 when fine */
static int check(const options* o)
{
  const char* what = NULL;
  const int from_header = !o->in && o->header;
  if (o->dims && o->n[0] * o->n[1] * o->n[2] * o->n[3] == 0)
    what = "array size must be nonzero";
  else if (!o->in && !o->zfp)
    what = "must specify uncompressed or compressed input file via -i or -z";
  else if (!zfp_type_size(o->type) && !from_header)
    what = o->in ? "must specify scalar type via -f, -d, or -t to compress"
                 : "must specify scalar type via -f, -d, or -t or header via -h to decompress";
  else if (!o->dims && !from_header)
    what = o->in ? "must specify array dimensions via -1, -2, -3, or -4 to compress"
                 : "must specify array dimensions via -1, -2, -3, or -4 or header via -h to decompress";
  else if (!o->mode && !from_header)
    what = o->in ? "must specify compression parameters via -a, -c, -p, or -r to compress"
                 : "must specify compression parameters via -a, -c, -p, or -r or header via -h to decompress";
  else if (o->stats && !o->in)
    what = "must specify input file via -i to compute stats";
  else if (!o->in && o->zfp && o->header && (zfp_type_size(o->type) || o->dims))
    what = "cannot specify both field type/size and header";
  if (what)
    fprintf(stderr, "%s\n", what);
  return what ? 1 : 0;
}

static FILE* open_in(const char* path) { return strcmp(path, "-") ? fopen(path, "rb") : stdin; }
static FILE* open_out(const char* path) { return strcmp(path, "-") ? fopen(path, "wb") : stdout; }
static void close_file(FILE* f)
{
  if (f != stdin && f != stdout)
    fclose(f);
}

/* whole file into a growing buffer; *size receives its length */
static void* slurp(FILE* f, size_t* size, size_t* capacity)
{
  size_t cap = 0x200, n = 0;
  unsigned char* buf = NULL;
  for (;;) {
    unsigned char* nb = (unsigned char*)realloc(buf, cap);
    if (!nb) {
      free(buf);
      return NULL;
    }
    buf = nb;
    n += fread(buf + n, 1, cap - n, f);
    if (n < cap)
      break;
    cap *= 2;
  }
  *size = n;
  *capacity = cap;
  return buf;
}

static void set_mode(zfp_stream* z, const options* o, zfp_field* f, int* ok)
{
  *ok = 1;
  switch (o->mode) {
    case 'R': zfp_stream_set_reversible(z); break;
    case 'a': zfp_stream_set_accuracy(z, o->tolerance); break;
    case 'p': zfp_stream_set_precision(z, o->precision); break;
    case 'r': zfp_stream_set_rate(z, o->rate, o->type, o->dims, zfp_false); break;
    case 'c':
      *ok = zfp_stream_set_params(z, o->minbits, o->maxbits ? o->maxbits : ZFP_MAX_BITS,
                                  o->maxprec ? o->maxprec : zfp_field_precision(f), o->minexp) != 0;
      break;
  }
}

static int set_policy(zfp_stream* z, const options* o)
{
  int ok = zfp_stream_set_execution(z, o->exec) != 0;
  if (ok && o->exec == zfp_exec_omp)
    ok = zfp_stream_set_omp_threads(z, o->threads) && zfp_stream_set_omp_chunk_size(z, o->chunk);
  if (!ok) {
    static const char* names[] = {"serial", "OpenMP", "cuda", "hip"};
    fprintf(stderr, "%s execution not available\n", names[o->exec]);
  }
  return ok;
}

/* value i of an array of `type` as a double */
static double value_at(const void* p, zfp_type type, size_t i)
{
  switch (type) {
    case zfp_type_int32: return (double)((const int32*)p)[i];
    case zfp_type_int64: return (double)((const int64*)p)[i];
    case zfp_type_float: return (double)((const float*)p)[i];
    default: return ((const double*)p)[i];
  }
}

/* rms, range-normalised rms, max error and PSNR of the reconstruction, in
   the reference tool's format (utils/zfp.c:30-80) */
static void print_stats(const void* a, const void* b, zfp_type type, size_t count)
{
  double lo = DBL_MAX, hi = -DBL_MAX, sum2 = 0, emax = 0;
  for (size_t i = 0; i < count; i++) {
    const double u = value_at(a, type, i);
    double e; /* the difference is taken in the element type, then widened */
    if (type == zfp_type_int32)
      e = fabs((double)(((const int32*)a)[i] - ((const int32*)b)[i]));
    else if (type == zfp_type_int64)
      e = fabs((double)(((const int64*)a)[i] - ((const int64*)b)[i]));
    else if (type == zfp_type_float)
      e = fabs((double)(((const float*)a)[i] - ((const float*)b)[i]));
    else
      e = fabs(u - value_at(b, type, i));
    if (e > emax) emax = e;
    sum2 += e * e;
    if (u < lo) lo = u;
    if (u > hi) hi = u;
  }
  const double rms = sqrt(sum2 / count);
  fprintf(stderr, " rmse=%.4g nrmse=%.4g maxe=%.4g psnr=%.2f", rms, rms / (hi - lo), emax,
          20 * log10((hi - lo) / (2 * rms)));
}

static int fail(const char* what)
{
  fprintf(stderr, "%s\n", what);
  return EXIT_FAILURE;
}

int main(int argc, char** argv)
{
  options o;
  parse(argc, argv, &o);
  if (check(&o))
    return EXIT_FAILURE;

  size_t tsize = zfp_type_size(o.type);
  size_t count = o.n[0] * o.n[1] * o.n[2] * o.n[3];
  size_t rawsize = 0, zfpsize = 0, cap = 0;
  void *raw = NULL, *rec = NULL, *buf = NULL;
  bitstream* bs = NULL;
  zfp_stream* z = zfp_stream_open(NULL);
  zfp_field* f = zfp_field_alloc();

  if (o.in) {
    FILE* fin = open_in(o.in);
    if (!fin)
      return fail("cannot open input file");
    rawsize = tsize * count;
    if (!(raw = malloc(rawsize)))
      return fail("cannot allocate memory");
    if (fread(raw, tsize, count, fin) != count)
      return fail("cannot read input file");
    close_file(fin);
    zfp_field_set_pointer(f, raw);
  }
  else {
    FILE* fz = open_in(o.zfp);
    if (!fz)
      return fail("cannot open compressed file");
    if (!(buf = slurp(fz, &zfpsize, &cap)))
      return fail("cannot allocate memory");
    if (ferror(fz))
      return fail("cannot read compressed file");
    close_file(fz);
    if (!(bs = stream_open(buf, cap)))
      return fail("cannot open compressed stream");
    zfp_stream_set_bit_stream(z, bs);
  }

  if (o.in || !o.header) {
    zfp_field_set_type(f, o.type);
    switch (o.dims) {
      case 1: zfp_field_set_size_1d(f, o.n[0]); break;
      case 2: zfp_field_set_size_2d(f, o.n[0], o.n[1]); break;
      case 3: zfp_field_set_size_3d(f, o.n[0], o.n[1], o.n[2]); break;
      case 4: zfp_field_set_size_4d(f, o.n[0], o.n[1], o.n[2], o.n[3]); break;
    }
    int ok;
    set_mode(z, &o, f, &ok);
    if (!ok)
      return fail("invalid compression parameters");
  }
  if (!set_policy(z, &o))
    return EXIT_FAILURE;

  if (o.in) {
    if (!(cap = zfp_stream_maximum_size(z, f)))
      return fail("invalid compression parameters");
    if (!(buf = malloc(cap)))
      return fail("cannot allocate memory");
    if (!(bs = stream_open(buf, cap)))
      return fail("cannot open compressed stream");
    zfp_stream_set_bit_stream(z, bs);
    if (o.header && !zfp_write_header(z, f, ZFP_HEADER_FULL))
      return fail("cannot write header");
    if (!(zfpsize = zfp_compress(z, f)))
      return fail("compression failed");
    if (o.zfp) {
      FILE* fz = open_out(o.zfp);
      if (!fz)
        return fail("cannot create compressed file");
      if (fwrite(buf, 1, zfpsize, fz) != zfpsize)
        return fail("cannot write compressed file");
      close_file(fz);
    }
  }

  if ((!o.in && o.zfp) || o.out || o.stats) {
    zfp_stream_rewind(z);
    if (o.header) {
      if (!zfp_read_header(z, f, ZFP_HEADER_FULL))
        return fail("incorrect or missing header");
      o.type = zfp_field_type(f);
      if (!(tsize = zfp_type_size(o.type)))
        return fail("unsupported type");
      o.n[0] = f->nx ? f->nx : 1;
      o.n[1] = f->ny ? f->ny : 1;
      o.n[2] = f->nz ? f->nz : 1;
      o.n[3] = f->nw ? f->nw : 1;
      count = o.n[0] * o.n[1] * o.n[2] * o.n[3];
    }
    rawsize = tsize * count;
    if (!(rec = malloc(rawsize)))
      return fail("cannot allocate memory");
    zfp_field_set_pointer(f, rec);
    if (!zfp_decompress(z, f))
      return fail("decompression failed");
    if (o.out) {
      FILE* fo = open_out(o.out);
      if (!fo)
        return fail("cannot create output file");
      if (fwrite(rec, tsize, count, fo) != count)
        return fail("cannot write output file");
      close_file(fo);
    }
  }

  if (!o.quiet) {
    static const char* tn[] = {"int32", "int64", "float", "double"};
    fprintf(stderr, "type=%s nx=%zu ny=%zu nz=%zu nw=%zu", tn[o.type - zfp_type_int32], o.n[0], o.n[1], o.n[2],
            o.n[3]);
    fprintf(stderr, " raw=%lu zfp=%lu ratio=%.3g rate=%.4g", (unsigned long)rawsize, (unsigned long)zfpsize,
            (double)rawsize / zfpsize, CHAR_BIT * (double)zfpsize / count);
    if (o.stats)
      print_stats(raw, rec, o.type, count);
    fprintf(stderr, "\n");
  }

  zfp_field_free(f);
  zfp_stream_close(z);
  stream_close(bs);
  free(buf);
  free(raw);
  free(rec);
  return EXIT_SUCCESS;
}
