// momentBasedD3Q19 -- D3Q19 solver driver (see momentBased.cpp).
int run_moment_based(int Q, int argc, char **argv);
int main(int argc, char **argv) { return run_moment_based(19, argc, argv); }
