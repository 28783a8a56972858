// The drop-in positions containers (localizer_amd/dropin/StorageHIP.h) on the input format of
// oracle/ref/ref_storage.cpp: writes the matrix and the text file for the byte-exact
// comparison with the reference's own containers (tests/test_storage_golden.py).
//   storage_main <in.bin> <matrix.bin> <text.txt> <header>
#include <cstdio>
#include <exception>

#include "../../localizer_amd/dropin/StorageHIP.h"
#include "../../oracle/ref/storage_io.h"

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: storage_main in.bin matrix.bin text.txt header\n");
        return 2;
    }
    try {
        return lzr_storage_io::run(argv[1], argv[2], argv[3], argv[4]);
    } catch (std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
