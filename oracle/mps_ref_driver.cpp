// mps_ref_driver.cpp -- a C entry point over the reference's own MPS reader
// (mps_reader::read_mps, cpp/mps_reader/include/mps_reader/mpsreader.hpp:19),
// compiled with the reader's sources where they lie by ref.mk into
// _ref/libmpsref.so.
//
// TEST INFRASTRUCTURE ONLY: it checks the MPS restatement (mps.py) and the
// engine's reader against the real reference on the same texts.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>
#include <vector>

#include "mps_reader/mpsreader.hpp"

extern "C" {

typedef struct mpsref_lp {
  int32_t m, n;
  int64_t nnz;
  int64_t* col_ptr;  // CSC: n + 1
  int32_t* row_idx;
  double* csc_val;
  int64_t* row_ptr;  // CSR: m + 1
  int32_t* col_idx;
  double* csr_val;
  double* c;
  double* AL;
  double* AU;
  double* l;
  double* u;
  double obj_constant;
  char* name;
} mpsref_lp;

void mpsref_free(mpsref_lp* lp);
int mpsref_read(const char* text, size_t length, mpsref_lp* out, char* error, size_t error_size);
}

namespace {

template <class T, class S>
T* dup(const std::vector<S>& v) {
  T* p = static_cast<T*>(std::malloc(sizeof(T) * (v.empty() ? 1 : v.size())));
  for (size_t k = 0; k < v.size(); ++k) p[k] = static_cast<T>(v[k]);
  return p;
}

}  // namespace

int mpsref_read(const char* text, size_t length, mpsref_lp* out, char* error, size_t error_size) {
  std::memset(out, 0, sizeof(*out));
  try {
    std::istringstream input(std::string(text, length));
    const mps_reader::MpsModel model = mps_reader::read_mps(input);
    const auto& lp = model.lp;
    const auto& csc = lp.csc();
    const auto& csr = lp.csr();
    out->m = lp.num_rows();
    out->n = lp.num_cols();
    out->nnz = static_cast<int64_t>(csc.values().size());
    out->col_ptr = dup<int64_t>(csc.col_ptr());
    out->row_idx = dup<int32_t>(csc.row_idx());
    out->csc_val = dup<double>(csc.values());
    out->row_ptr = dup<int64_t>(csr.row_ptr());
    out->col_idx = dup<int32_t>(csr.col_idx());
    out->csr_val = dup<double>(csr.values());
    out->c = dup<double>(lp.objective());
    out->AL = dup<double>(lp.row_lower());
    out->AU = dup<double>(lp.row_upper());
    out->l = dup<double>(lp.col_lower());
    out->u = dup<double>(lp.col_upper());
    out->obj_constant = lp.obj_constant();
    out->name = static_cast<char*>(std::malloc(model.name.size() + 1));
    std::memcpy(out->name, model.name.c_str(), model.name.size() + 1);
    return 0;
  } catch (const std::exception& e) {
    mpsref_free(out);
    if (error && error_size) {
      std::strncpy(error, e.what(), error_size - 1);
      error[error_size - 1] = '\0';
    }
    return 1;
  }
}

void mpsref_free(mpsref_lp* lp) {
  if (!lp) return;
  void* arrays[] = {lp->col_ptr, lp->row_idx, lp->csc_val, lp->row_ptr, lp->col_idx, lp->csr_val, lp->c,
                    lp->AL,      lp->AU,      lp->l,       lp->u,       lp->name};
  for (void* p : arrays) std::free(p);
  std::memset(lp, 0, sizeof(*lp));
}
