/* placeholder replaced below */
double or_cpu_rate_f32(int nx, int ny, int nz, int pml, int steps, int repeats, int threads) { (void)nx;(void)ny;(void)nz;(void)pml;(void)steps;(void)repeats;(void)threads; return 0; }
