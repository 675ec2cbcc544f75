// analyze_cmd.cpp -- `kite analyze` (placeholder until the correlation engine lands).
#include "cli.hpp"
#include "commands.hpp"
#include "kitefmt.hpp"

namespace kh {
int run_analyze(int argc, char **argv) {
    (void)argc;
    (void)argv;
    throw Error("analyze is not built yet");
}
}  // namespace kh
