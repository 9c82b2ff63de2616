// Matrix ingest: sparse matrices from files (Matrix Market coordinate,
// Harwell-Boeing assembled real/pattern, plain coordinate triples) into
// coordinate form on the host, and coordinate -> CSR on the device (sort by
// (row, column), duplicates summed in input order), feeding the device
// operator.  These are the formats the reference's example drivers read
// (README:103-118: Harwell-Boeing, coordinate; Matrix Market for further
// test matrices).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pk_internal.h"

namespace pk {
void set_error(const char* fmt, ...);

namespace {

// ------------------------------------------------------------ host parse --
struct Coo {
  int m = 0, n = 0;
  std::vector<int> r, c;  // 0-based
  std::vector<double> v;
  bool cplx = false;      // a complex Matrix Market file: imaginary parts in vi
  std::vector<double> vi;
};

std::string lower(std::string s) {
  for (char& ch : s) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
  return s;
}

bool read_lines(const char* path, std::vector<std::string>& lines) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::string cur;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) {
    for (size_t i = 0; i < got; ++i) {
      if (buf[i] == '\n') {
        if (!cur.empty() && cur.back() == '\r') cur.pop_back();
        lines.push_back(cur);
        cur.clear();
      } else {
        cur.push_back(buf[i]);
      }
    }
  }
  if (!cur.empty()) lines.push_back(cur);
  std::fclose(f);
  return true;
}

// Fortran-style real: 'D'/'d' exponents accepted.
double parse_real(std::string s) {
  for (char& ch : s)
    if (ch == 'D' || ch == 'd') ch = 'E';
  return std::strtod(s.c_str(), nullptr);
}

// Symmetric storage: add the mirrored entries (kind 1 symmetric, 2 skew:
// negated, 3 Hermitian: conjugated).
void expand_symmetry(Coo& A, int kind) {
  if (kind == 0) return;
  const size_t nz = A.r.size();
  for (size_t q = 0; q < nz; ++q) {
    if (A.r[q] == A.c[q]) continue;
    A.r.push_back(A.c[q]);
    A.c.push_back(A.r[q]);
    A.v.push_back(kind == 2 ? -A.v[q] : A.v[q]);
    if (A.cplx) A.vi.push_back(kind == 2 ? -A.vi[q] : kind == 3 ? -A.vi[q] : A.vi[q]);
  }
}

// Matrix Market coordinate: real / integer / pattern / complex, general /
// symmetric / skew-symmetric / Hermitian.
int parse_mm(const std::vector<std::string>& L, Coo& A) {
  const std::string h = lower(L[0]);
  if (h.find("coordinate") == std::string::npos) {
    set_error("Matrix Market: only the coordinate format is read");
    return PK_ERR_NOTSUP;
  }
  A.cplx = h.find("complex") != std::string::npos;
  const bool pattern = h.find("pattern") != std::string::npos;
  const int sym = h.find("skew-symmetric") != std::string::npos ? 2
                  : h.find("hermitian") != std::string::npos  ? (A.cplx ? 3 : 1)
                  : h.find("symmetric") != std::string::npos  ? 1
                                                              : 0;
  size_t i = 1;
  while (i < L.size() && (L[i].empty() || L[i][0] == '%')) ++i;
  if (i >= L.size()) return PK_ERR_ARG;
  long long nz = 0;
  if (std::sscanf(L[i].c_str(), "%d %d %lld", &A.m, &A.n, &nz) != 3 || A.m < 0 || A.n < 0 ||
      nz < 0) {
    set_error("Matrix Market: bad size line");
    return PK_ERR_ARG;
  }
  A.r.reserve(nz);
  A.c.reserve(nz);
  A.v.reserve(nz);
  if (A.cplx) A.vi.reserve(nz);
  for (++i; i < L.size() && static_cast<long long>(A.r.size()) < nz; ++i) {
    if (L[i].empty() || L[i][0] == '%') continue;
    const char* p = L[i].c_str();
    char* e;
    char* s;
    bool ok = true;
    const long ri = std::strtol(p, &e, 10);
    ok = ok && e != p;
    const long ci = std::strtol(s = e, &e, 10);
    ok = ok && e != s;
    double v = 1.0;
    if (!pattern) {
      v = std::strtod(s = e, &e);
      ok = ok && e != s;
    }
    if (A.cplx) {
      A.vi.push_back(std::strtod(s = e, &e));
      ok = ok && e != s;
    }
    if (!ok) {
      set_error("Matrix Market: unreadable entry on line %zu", i + 1);
      return PK_ERR_ARG;
    }
    if (ri < 1 || ri > A.m || ci < 1 || ci > A.n) {
      set_error("Matrix Market: entry (%ld, %ld) outside %d x %d", ri, ci, A.m, A.n);
      return PK_ERR_ARG;
    }
    A.r.push_back(static_cast<int>(ri - 1));
    A.c.push_back(static_cast<int>(ci - 1));
    A.v.push_back(v);
  }
  if (static_cast<long long>(A.r.size()) != nz) {
    set_error("Matrix Market: %lld entries announced, %zu read", nz, A.r.size());
    return PK_ERR_ARG;
  }
  expand_symmetry(A, sym);
  return PK_OK;
}

// Fortran edit descriptor of a Harwell-Boeing header, e.g. (10I8),
// (1P,4E20.12), (3E25.16): fields per line and field width.
bool hb_format(std::string f, int* per, int* width) {
  f = lower(f);
  const size_t a = f.find('('), b = f.rfind(')');
  if (a == std::string::npos || b == std::string::npos || b <= a) return false;
  std::string s = f.substr(a + 1, b - a - 1);
  const size_t comma = s.rfind(',');  // drop a scale factor such as 1p,
  if (comma != std::string::npos) s = s.substr(comma + 1);
  size_t k = 0;
  while (k < s.size() && std::isdigit(static_cast<unsigned char>(s[k]))) ++k;
  *per = k ? std::atoi(s.substr(0, k).c_str()) : 1;
  while (k < s.size() && std::isalpha(static_cast<unsigned char>(s[k]))) ++k;
  *width = std::atoi(s.c_str() + k);
  return *per > 0 && *width > 0;
}

// count fixed-width fields from lines [i, ...) into out (as strings)
bool hb_fields(const std::vector<std::string>& L, size_t& i, long long count, int per, int width,
               std::vector<std::string>& out) {
  out.clear();
  out.reserve(count);
  while (static_cast<long long>(out.size()) < count) {
    if (i >= L.size()) return false;
    const std::string& s = L[i++];
    for (int f = 0; f < per && static_cast<long long>(out.size()) < count; ++f) {
      const size_t p = static_cast<size_t>(f) * width;
      if (p >= s.size()) break;
      out.push_back(s.substr(p, width));
    }
  }
  return true;
}

// Harwell-Boeing, assembled real or pattern matrices (RUA, RRA, RSA, RZA,
// PUA, PRA, PSA): column pointers, row indices, values.
int parse_hb(const std::vector<std::string>& L, Coo& A) {
  if (L.size() < 5) return PK_ERR_ARG;
  long long tot = 0, ptr = 0, ind = 0, val = 0, rhs = 0;
  std::sscanf(L[1].c_str(), "%lld %lld %lld %lld %lld", &tot, &ptr, &ind, &val, &rhs);
  const std::string type = lower(L[2].substr(0, 3));
  long long nrow = 0, ncol = 0, nnz = 0;
  if (L[2].size() < 4 ||
      std::sscanf(L[2].c_str() + 3, "%lld %lld %lld", &nrow, &ncol, &nnz) != 3) {
    set_error("Harwell-Boeing: bad type line");
    return PK_ERR_ARG;
  }
  if (type.size() < 3 || (type[0] != 'r' && type[0] != 'p') || type[2] != 'a') {
    set_error("Harwell-Boeing: type %s is not an assembled real or pattern matrix",
              type.c_str());
    return PK_ERR_NOTSUP;
  }
  const std::string& fl = L[3];
  int pp, pw, ip, iw, vp = 1, vw = 1;
  if (!hb_format(fl.substr(0, std::min<size_t>(16, fl.size())), &pp, &pw) ||
      !hb_format(fl.substr(16, 16), &ip, &iw) ||
      (type[0] == 'r' && !hb_format(fl.substr(32, 20), &vp, &vw))) {
    set_error("Harwell-Boeing: unreadable formats line");
    return PK_ERR_ARG;
  }
  size_t i = rhs > 0 ? 5 : 4;
  std::vector<std::string> fp, fi, fv;
  if (!hb_fields(L, i, ncol + 1, pp, pw, fp) || !hb_fields(L, i, nnz, ip, iw, fi) ||
      (type[0] == 'r' && !hb_fields(L, i, nnz, vp, vw, fv))) {
    set_error("Harwell-Boeing: file ends early");
    return PK_ERR_ARG;
  }
  A.m = static_cast<int>(nrow);
  A.n = static_cast<int>(ncol);
  A.r.resize(nnz);
  A.c.resize(nnz);
  A.v.resize(nnz);
  for (long long j = 0; j < ncol; ++j) {
    const long long a = std::atoll(fp[j].c_str()) - 1, b = std::atoll(fp[j + 1].c_str()) - 1;
    if (a < 0 || b < a || b > nnz) {
      set_error("Harwell-Boeing: bad column pointer");
      return PK_ERR_ARG;
    }
    for (long long q = a; q < b; ++q) {
      const long long r = std::atoll(fi[q].c_str());
      if (r < 1 || r > nrow) {
        set_error("Harwell-Boeing: row index %lld outside 1..%lld", r, nrow);
        return PK_ERR_ARG;
      }
      A.r[q] = static_cast<int>(r - 1);
      A.c[q] = static_cast<int>(j);
      A.v[q] = type[0] == 'r' ? parse_real(fv[q]) : 1.0;
    }
  }
  expand_symmetry(A, type[1] == 's' ? 1 : (type[1] == 'z' ? 2 : 0));
  return PK_OK;
}

// Plain coordinate triples "i j value" (1-based), optionally after a line
// "m n nnz" announcing the size and the entry count.
int parse_coord(const std::vector<std::string>& L, Coo& A) {
  std::vector<size_t> data;
  for (size_t i = 0; i < L.size(); ++i) {
    const std::string& s = L[i];
    size_t k = 0;
    while (k < s.size() && std::isspace(static_cast<unsigned char>(s[k]))) ++k;
    if (k == s.size() || s[k] == '%' || s[k] == '#') continue;
    data.push_back(i);
  }
  size_t first = 0;
  long long hm = -1, hn = -1, hz = -1;
  if (!data.empty()) {
    char tail[8];
    if (std::sscanf(L[data[0]].c_str(), "%lld %lld %lld %7s", &hm, &hn, &hz, tail) == 3 &&
        hz == static_cast<long long>(data.size()) - 1 && L[data[0]].find('.') == std::string::npos &&
        L[data[0]].find('e') == std::string::npos && L[data[0]].find('E') == std::string::npos)
      first = 1;
    else
      hm = hn = -1;
  }
  int mm = 0, nn = 0;
  for (size_t t = first; t < data.size(); ++t) {
    const char* p = L[data[t]].c_str();
    char* e;
    const long ri = std::strtol(p, &e, 10);
    const long ci = std::strtol(e, &e, 10);
    const double v = std::strtod(e, nullptr);
    if (ri < 1 || ci < 1) {
      set_error("coordinate file: bad entry on line %zu", data[t] + 1);
      return PK_ERR_ARG;
    }
    A.r.push_back(static_cast<int>(ri - 1));
    A.c.push_back(static_cast<int>(ci - 1));
    A.v.push_back(v);
    mm = std::max(mm, static_cast<int>(ri));
    nn = std::max(nn, static_cast<int>(ci));
  }
  A.m = hm >= 0 ? static_cast<int>(hm) : mm;
  A.n = hn >= 0 ? static_cast<int>(hn) : nn;
  if (mm > A.m || nn > A.n) {
    set_error("coordinate file: entries outside the announced %d x %d", A.m, A.n);
    return PK_ERR_ARG;
  }
  return PK_OK;
}

int read_any(const char* path, Coo& A) {
  std::vector<std::string> L;
  if (!path || !read_lines(path, L) || L.empty()) {
    set_error("cannot read %s", path ? path : "(null)");
    return PK_ERR_ARG;
  }
  if (L[0].rfind("%%MatrixMarket", 0) == 0) return parse_mm(L, A);
  if (L.size() >= 4) {
    const std::string t = L.size() > 2 && L[2].size() >= 3 ? lower(L[2].substr(0, 3)) : "";
    long long a, b, c, d;
    if (t.size() == 3 && std::isalpha(static_cast<unsigned char>(t[0])) &&
        std::isalpha(static_cast<unsigned char>(t[1])) &&
        std::isalpha(static_cast<unsigned char>(t[2])) &&
        std::sscanf(L[1].c_str(), "%lld %lld %lld %lld", &a, &b, &c, &d) == 4 &&
        L[3].find('(') != std::string::npos)
      return parse_hb(L, A);
  }
  return parse_coord(L, A);
}

// ---------------------------------------------------------- device build --
__global__ void k_coo_keys(long long nnz, const int* __restrict__ r, const int* __restrict__ c,
                           int base, int n, unsigned long long* __restrict__ key,
                           int* __restrict__ perm) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    key[q] = static_cast<unsigned long long>(r[q] - base) * static_cast<unsigned long long>(n) +
             static_cast<unsigned long long>(c[q] - base);
    perm[q] = static_cast<int>(q);
  }
}

__global__ void k_coo_check(long long nnz, const int* __restrict__ r, const int* __restrict__ c,
                            int base, int m, int n, int* __restrict__ bad) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int ri = r[q] - base, ci = c[q] - base;
    if (ri < 0 || ri >= m || ci < 0 || ci >= n) bad[0] = 1;
  }
}

__global__ void k_heads(long long nnz, const unsigned long long* __restrict__ key,
                        int* __restrict__ head) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz;
       q += static_cast<long long>(gridDim.x) * blockDim.x)
    head[q] = (q == 0 || key[q] != key[q - 1]) ? 1 : 0;
}

// One thread per run of equal keys: its values summed in input order (the
// sort is stable), column and row of the entry.
template <typename VT>
__global__ void k_sum_runs(long long nnz, const unsigned long long* __restrict__ key,
                           const int* __restrict__ perm, const int* __restrict__ head,
                           const int* __restrict__ slot, const VT* __restrict__ val, int n,
                           int* __restrict__ urow, int* __restrict__ ucol,
                           VT* __restrict__ uval) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (!head[q]) continue;
    VT s = val[perm[q]];
    for (long long t = q + 1; t < nnz && !head[t]; ++t) s += val[perm[t]];
    const int u = slot[q];
    uval[u] = s;
    ucol[u] = static_cast<int>(key[q] % static_cast<unsigned long long>(n));
    urow[u] = static_cast<int>(key[q] / static_cast<unsigned long long>(n));
  }
}

__global__ void k_rowptr_rows(long long nz, int m, const int* __restrict__ rows,
                              int* __restrict__ rp) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q <= nz;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int lo = q == 0 ? -1 : rows[q - 1];
    const int hi = q == nz ? m : rows[q];
    for (int j = lo + 1; j <= hi; ++j) rp[j] = static_cast<int>(q);
  }
}

int bits64(unsigned long long v) {
  int b = 1;
  while (b < 64 && (1ULL << b) < v) ++b;
  return b;
}

// COO (host or device arrays, base 0 or 1) -> device CSR (rp m+1, ci/va
// nz_out entries, allocated here).
template <typename VT>
int coo_to_csr_dev(int m, int n, long long nnz, const int* rows, const int* cols, const VT* vals,
                   int where, int base, int** rp_out, int** ci_out, VT** va_out,
                   long long* nz_out, hipStream_t s) {
  *rp_out = nullptr;
  *ci_out = nullptr;
  *va_out = nullptr;
  int rc = PK_OK;
  const size_t nz1 = static_cast<size_t>(std::max<long long>(nnz, 1));
  int *dr = nullptr, *dc = nullptr, *perm_in = nullptr, *perm_out = nullptr, *head = nullptr,
      *slot = nullptr, *urow = nullptr, *ucol = nullptr, *rp = nullptr, *bad = nullptr;
  VT *dv = nullptr, *uval = nullptr;
  unsigned long long *kin = nullptr, *kout = nullptr;
  void* tmp = nullptr;
  size_t tb = 0, tb2 = 0;
  int hbad = 0, nuniq = 0, last_head = 0;
  const int grid = 2048;
#define PKI(x)             \
  if ((x) != hipSuccess) { \
    rc = PK_ERR_HIP;       \
    goto out;              \
  }
  if (where == PK_HOST) {
    PKI(hipMalloc(&dr, sizeof(int) * nz1));
    PKI(hipMalloc(&dc, sizeof(int) * nz1));
    PKI(hipMalloc(&dv, sizeof(VT) * nz1));
    if (nnz > 0) {
      PKI(hipMemcpyAsync(dr, rows, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
      PKI(hipMemcpyAsync(dc, cols, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
      PKI(hipMemcpyAsync(dv, vals, sizeof(VT) * nnz, hipMemcpyHostToDevice, s));
    }
  }
  {
    const int* R = where == PK_HOST ? dr : rows;
    const int* Cc = where == PK_HOST ? dc : cols;
    const VT* V = where == PK_HOST ? dv : vals;
    PKI(hipMalloc(&bad, sizeof(int)));
    PKI(hipMemsetAsync(bad, 0, sizeof(int), s));
    PKI(hipMalloc(&kin, sizeof(unsigned long long) * nz1));
    PKI(hipMalloc(&kout, sizeof(unsigned long long) * nz1));
    PKI(hipMalloc(&perm_in, sizeof(int) * nz1));
    PKI(hipMalloc(&perm_out, sizeof(int) * nz1));
    PKI(hipMalloc(&head, sizeof(int) * nz1));
    PKI(hipMalloc(&slot, sizeof(int) * nz1));
    PKI(hipMalloc(&rp, sizeof(int) * (static_cast<size_t>(m) + 1)));
    if (nnz > 0) {
      hipLaunchKernelGGL(k_coo_check, dim3(grid), dim3(256), 0, s, nnz, R, Cc, base, m, n, bad);
      PKI(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
      PKI(hipStreamSynchronize(s));
      if (hbad) {
        set_error("coordinate entries outside %d x %d (base %d)", m, n, base);
        rc = PK_ERR_ARG;
        goto out;
      }
      hipLaunchKernelGGL(k_coo_keys, dim3(grid), dim3(256), 0, s, nnz, R, Cc, base, n, kin,
                         perm_in);
      const int nbits = bits64(static_cast<unsigned long long>(m) * static_cast<unsigned long long>(n));
      PKI(rocprim::radix_sort_pairs(nullptr, tb, kin, kout, perm_in, perm_out,
                                    static_cast<size_t>(nnz), 0, nbits, s));
      PKI(rocprim::exclusive_scan(nullptr, tb2, head, slot, 0, static_cast<size_t>(nnz),
                                  rocprim::plus<int>(), s));
      PKI(hipMalloc(&tmp, std::max(tb, tb2)));
      PKI(rocprim::radix_sort_pairs(tmp, tb, kin, kout, perm_in, perm_out,
                                    static_cast<size_t>(nnz), 0, nbits, s));
      hipLaunchKernelGGL(k_heads, dim3(grid), dim3(256), 0, s, nnz, kout, head);
      PKI(rocprim::exclusive_scan(tmp, tb2, head, slot, 0, static_cast<size_t>(nnz),
                                  rocprim::plus<int>(), s));
      PKI(hipMemcpyAsync(&nuniq, slot + (nnz - 1), sizeof(int), hipMemcpyDeviceToHost, s));
      PKI(hipMemcpyAsync(&last_head, head + (nnz - 1), sizeof(int), hipMemcpyDeviceToHost, s));
      PKI(hipStreamSynchronize(s));
      nuniq += last_head;
    }
    const size_t nu1 = static_cast<size_t>(std::max(nuniq, 1));
    PKI(hipMalloc(&urow, sizeof(int) * nu1));
    PKI(hipMalloc(&ucol, sizeof(int) * nu1));
    PKI(hipMalloc(&uval, sizeof(VT) * nu1));
    if (nnz > 0)
      hipLaunchKernelGGL((k_sum_runs<VT>), dim3(grid), dim3(256), 0, s, nnz, kout, perm_out, head,
                         slot, V, n, urow, ucol, uval);
    hipLaunchKernelGGL(k_rowptr_rows, dim3(grid), dim3(256), 0, s, static_cast<long long>(nuniq),
                       m, urow, rp);
    PKI(hipGetLastError());
    PKI(hipStreamSynchronize(s));
  }
  *rp_out = rp;
  *ci_out = ucol;
  *va_out = uval;
  *nz_out = nuniq;
  rp = nullptr;
  ucol = nullptr;
  uval = nullptr;
out:
#undef PKI
  (void)hipFree(dr);
  (void)hipFree(dc);
  (void)hipFree(dv);
  (void)hipFree(bad);
  (void)hipFree(kin);
  (void)hipFree(kout);
  (void)hipFree(perm_in);
  (void)hipFree(perm_out);
  (void)hipFree(head);
  (void)hipFree(slot);
  (void)hipFree(urow);
  (void)hipFree(ucol);
  (void)hipFree(uval);
  (void)hipFree(rp);
  (void)hipFree(tmp);
  return rc;
}

template <typename VT>
int coo_to_csr_any(int m, int n, long long nnz, const int* rows, const int* cols, const void* vals,
                   int where, int base, int* rowptr, int* colind, void* val_out,
                   long long* nnz_out) {
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PK_ERR_HIP;
  int *rp = nullptr, *ci = nullptr;
  VT* va = nullptr;
  long long nz = 0;
  int rc = coo_to_csr_dev<VT>(m, n, nnz, rows, cols, static_cast<const VT*>(vals), where, base,
                              &rp, &ci, &va, &nz, s);
  if (rc == PK_OK) {
    const hipMemcpyKind k = where == PK_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (hipMemcpy(rowptr, rp, sizeof(int) * (static_cast<size_t>(m) + 1), k) != hipSuccess ||
        (nz > 0 && (hipMemcpy(colind, ci, sizeof(int) * nz, k) != hipSuccess ||
                    hipMemcpy(val_out, va, sizeof(VT) * nz, k) != hipSuccess)))
      rc = PK_ERR_HIP;
    *nnz_out = nz;
  }
  (void)hipFree(rp);
  (void)hipFree(ci);
  (void)hipFree(va);
  (void)hipStreamDestroy(s);
  return rc;
}

template <typename VT>
int create_from_coo(int m, int n, long long nnz, const int* rows, const int* cols,
                    const void* vals, int where, int base, int val_prec, CsrOp** op) {
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PK_ERR_HIP;
  int *rp = nullptr, *ci = nullptr;
  VT* va = nullptr;
  long long nz = 0;
  int rc = coo_to_csr_dev<VT>(m, n, nnz, rows, cols, static_cast<const VT*>(vals), where, base,
                              &rp, &ci, &va, &nz, s);
  if (rc == PK_OK) rc = csr_create(m, n, nz, rp, ci, va, val_prec, PK_DEVICE, op);
  (void)hipFree(rp);
  (void)hipFree(ci);
  (void)hipFree(va);
  (void)hipStreamDestroy(s);
  return rc;
}

bool base_ok(int base) { return base == 0 || base == 1; }

}  // namespace
}  // namespace pk

using namespace pk;

extern "C" {

int pk_matrix_read(const char* path, int* m, int* n, long long* nnz, int** rows, int** cols,
                   double** vals) {
  if (!m || !n || !nnz || !rows || !cols || !vals) return PK_ERR_ARG;
  Coo A;
  const int rc = read_any(path, A);
  if (rc != PK_OK) return rc;
  if (A.cplx) {
    set_error("pk_matrix_read: %s holds a complex matrix (pk_csr_read with PK_C128 or PK_C64 "
              "reads it)", path);
    return PK_ERR_NOTSUP;
  }
  const size_t nz = A.r.size();
  *rows = static_cast<int*>(std::malloc(sizeof(int) * std::max<size_t>(nz, 1)));
  *cols = static_cast<int*>(std::malloc(sizeof(int) * std::max<size_t>(nz, 1)));
  *vals = static_cast<double*>(std::malloc(sizeof(double) * std::max<size_t>(nz, 1)));
  if (!*rows || !*cols || !*vals) {
    std::free(*rows);
    std::free(*cols);
    std::free(*vals);
    return PK_ERR_OOM;
  }
  if (nz) {
    std::memcpy(*rows, A.r.data(), sizeof(int) * nz);
    std::memcpy(*cols, A.c.data(), sizeof(int) * nz);
    std::memcpy(*vals, A.v.data(), sizeof(double) * nz);
  }
  *m = A.m;
  *n = A.n;
  *nnz = static_cast<long long>(nz);
  return PK_OK;
}

void pk_matrix_free(void* p) { std::free(p); }

int pk_coo_to_csr(int m, int n, long long nnz, const int* rows, const int* cols, const void* vals,
                  int val_prec, int where, int base, int* rowptr, int* colind, void* val_out,
                  long long* nnz_out) {
  if (m < 0 || n < 0 || nnz < 0 || nnz >= (1LL << 31) || !base_ok(base) || !rowptr ||
      !nnz_out || (nnz > 0 && (!rows || !cols || !vals || !colind || !val_out)) ||
      (where != PK_HOST && where != PK_DEVICE))
    return PK_ERR_ARG;
  if (val_prec == PK_F64)
    return coo_to_csr_any<double>(m, n, nnz, rows, cols, vals, where, base, rowptr, colind,
                                  val_out, nnz_out);
  if (val_prec == PK_F32)
    return coo_to_csr_any<float>(m, n, nnz, rows, cols, vals, where, base, rowptr, colind,
                                 val_out, nnz_out);
  if (val_prec == PK_C128)
    return coo_to_csr_any<double2>(m, n, nnz, rows, cols, vals, where, base, rowptr, colind,
                                   val_out, nnz_out);
  if (val_prec == PK_C64)
    return coo_to_csr_any<float2>(m, n, nnz, rows, cols, vals, where, base, rowptr, colind,
                                  val_out, nnz_out);
  return PK_ERR_ARG;
}

int pk_csr_create_coo(int m, int n, long long nnz, const int* rows, const int* cols,
                      const void* vals, int val_prec, int where, int base, pk_csr_t* op) {
  if (!op || m < 0 || n < 0 || nnz < 0 || nnz >= (1LL << 31) || !base_ok(base) ||
      (nnz > 0 && (!rows || !cols || !vals)) || (where != PK_HOST && where != PK_DEVICE))
    return PK_ERR_ARG;
  CsrOp* o = nullptr;
  int rc;
  if (val_prec == PK_F64)
    rc = create_from_coo<double>(m, n, nnz, rows, cols, vals, where, base, val_prec, &o);
  else if (val_prec == PK_F32)
    rc = create_from_coo<float>(m, n, nnz, rows, cols, vals, where, base, val_prec, &o);
  else if (val_prec == PK_C128)
    rc = create_from_coo<double2>(m, n, nnz, rows, cols, vals, where, base, val_prec, &o);
  else if (val_prec == PK_C64)
    rc = create_from_coo<float2>(m, n, nnz, rows, cols, vals, where, base, val_prec, &o);
  else
    return PK_ERR_ARG;
  if (rc == PK_OK) *op = reinterpret_cast<pk_csr_t>(o);
  return rc;
}

int pk_csr_read(const char* path, int val_prec, pk_csr_t* op) {
  if (!op || (val_prec != PK_F64 && val_prec != PK_F32 && !prec_is_complex(val_prec)))
    return PK_ERR_ARG;
  Coo A;
  int rc = read_any(path, A);
  if (rc != PK_OK) return rc;
  const long long nz = static_cast<long long>(A.r.size());
  if (prec_is_complex(val_prec)) {
    // (re, im) pairs; a real file reads with zero imaginary parts
    std::vector<double> z(2 * static_cast<size_t>(nz));
    for (long long q = 0; q < nz; ++q) {
      z[2 * q] = A.v[q];
      z[2 * q + 1] = A.cplx ? A.vi[q] : 0.0;
    }
    if (val_prec == PK_C128)
      return pk_csr_create_coo(A.m, A.n, nz, A.r.data(), A.c.data(), z.data(), PK_C128, PK_HOST,
                               0, op);
    std::vector<float> z32(z.begin(), z.end());
    return pk_csr_create_coo(A.m, A.n, nz, A.r.data(), A.c.data(), z32.data(), PK_C64, PK_HOST,
                             0, op);
  }
  if (A.cplx) {
    set_error("pk_csr_read: %s holds a complex matrix (read it with PK_C128 or PK_C64)", path);
    return PK_ERR_NOTSUP;
  }
  if (val_prec == PK_F64)
    return pk_csr_create_coo(A.m, A.n, nz, A.r.data(), A.c.data(), A.v.data(), PK_F64, PK_HOST,
                             0, op);
  std::vector<float> v32(A.v.begin(), A.v.end());
  return pk_csr_create_coo(A.m, A.n, nz, A.r.data(), A.c.data(), v32.data(), PK_F32, PK_HOST, 0,
                           op);
}

}  // extern "C"
