// commands.hpp -- the `kite` subcommands implemented by the host.
#pragma once

#include <string>

namespace kh {

int run_count(int argc, char **argv);
int run_dump(int argc, char **argv);
int run_histo(int argc, char **argv);
int run_merge(int argc, char **argv);
int run_info(int argc, char **argv);
int run_analyze(int argc, char **argv);
int run_profile(int argc, char **argv);

// "[kite <cmd>] Command: ..." echo (src/kiteops/mod.rs:72-97).
void print_command_line(const char *cmd, int argc, char **argv);
// create_dir_all of the output's parent directory.
void make_parent_dirs(const std::string &path);
unsigned worker_threads(unsigned requested);

}  // namespace kh
