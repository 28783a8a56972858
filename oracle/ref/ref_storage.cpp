// Harness over the reference's positions containers (source/Storage.cpp, compiled where it
// lies): reads records from a binary file, fills the container of the given type, writes
// getLocalizedPositionsAsMatrix (column-major doubles) and writePositionsToFile's text.
// TEST INFRASTRUCTURE ONLY (oracle/ref/Makefile builds it into oracle/_ref/).
//   ref_storage <in.bin> <matrix.bin> <text.txt> <header>
// in.bin: int32 type (0 symmetric, 1 fixed width, 5 ellipsoidal), int32 n, then per record
// uint64 frameNumber, uint64 nFramesPresent and the type's doubles in declaration order
// (source/Storage.h:117-186).
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "Storage.h"

#include "storage_io.h"

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: ref_storage in.bin matrix.bin text.txt header\n");
        return 2;
    }
    try {
        return lzr_storage_io::run(argv[1], argv[2], argv[3], argv[4]);
    } catch (std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
