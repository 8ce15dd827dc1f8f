// pma: the reference's console program (Main.cpp, CommandLineDispatcher.cpp:27-150) over the C ABI.
//
//   pma -all|-edt|-rib|-apt|-spd [base folder]
//
// Runs the mode on every folder under the base folder (default: the current directory) that
// holds config.txt (Path::FillFoldersRecursively), in name order, on GPU 0.  Unlike the
// reference, which printed errors and exited 0, a failure ends the program with exit status 1.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "pma/pma.h"

namespace fs = std::filesystem;

int main(int argc, char** argv) {
    struct M {
        const char* flag;
        pma_mode mode;
    };
    const M modes[] = {{"-all", PMA_MODE_ALL}, {"-edt", PMA_MODE_EDT}, {"-rib", PMA_MODE_RIB},
                       {"-apt", PMA_MODE_APT}, {"-spd", PMA_MODE_SPD}};
    const M* m = nullptr;
    for (const M& x : modes)
        if (argc >= 2 && std::strcmp(argv[1], x.flag) == 0) m = &x;
    if (!m) {
        std::printf(
            "Console parameters are incorrect. Exiting.\n\n"
            "The following console parameters are supported:\n"
            "-edt - calculates Euclidean distance transform (maximum inscribed balls)\n"
            "-rib - removes inner balls (balls fully contained in other balls)\n"
            "-apt - assigns pores and throats\n"
            "-all - determines maximum inscribed balls, medial axis, pores, throats, and pore connectivity statistics\n"
            "-spd - computes shortest path distances (requires only initial binary image)\n");
        return 2;
    }
    const fs::path base = argc >= 3 ? fs::path(argv[2]) : fs::current_path();
    std::vector<fs::path> folders;
    std::error_code ec;
    if (fs::exists(base / "config.txt")) folders.push_back(base);
    for (auto it = fs::recursive_directory_iterator(base, ec); !ec && it != fs::recursive_directory_iterator(); it.increment(ec))
        if (it->is_directory() && fs::exists(it->path() / "config.txt")) folders.push_back(it->path());
    std::sort(folders.begin(), folders.end());
    if (folders.empty()) {
        std::fprintf(stderr, "pma: no folder with config.txt under %s\n", base.string().c_str());
        return 1;
    }
    pma_ctx* ctx = nullptr;
    if (pma_create(&ctx, 0) != PMA_OK) {
        std::fprintf(stderr, "pma: %s\n", pma_last_error());
        return 1;
    }
    int rc = 0;
    for (const fs::path& f : folders) {
        std::printf("%s %s\n", m->flag, f.string().c_str());
        std::fflush(stdout);
        if (pma_run_folder(ctx, f.string().c_str(), m->mode) != PMA_OK) {
            std::fprintf(stderr, "pma: %s: %s\n", f.string().c_str(), pma_last_error());
            rc = 1;
            break;
        }
    }
    pma_destroy(ctx);
    return rc;
}
