// momentBasedD3Q27 -- D3Q27 solver driver (see momentBased.cpp).
int run_moment_based(int Q, int argc, char **argv);
int main(int argc, char **argv) { return run_moment_based(27, argc, argv); }
