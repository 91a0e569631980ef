// The library's error sink (defined with the C ABI in d4_api.cpp), for a
// build of d4_ledger.cpp on its own.
#include <cstdio>
#include <string>
namespace d4 {
void set_error(const std::string &s) { fprintf(stderr, "  [error] %s\n", s.c_str()); }
}
