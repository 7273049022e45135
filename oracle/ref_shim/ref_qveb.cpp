// Binds the reference's classic vEB table (cpu/src/query-veb.cpp:25) so the
// hand-checkable small tables can be compared.
#include "query-veb.h"
extern "C" void ref_build_table_classic(void *table, uint64_t n, uint32_t d) {
    buildTable((vEB_table *)table, n, d, 0);
}
