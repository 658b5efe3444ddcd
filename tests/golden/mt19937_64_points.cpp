// Writes tests/golden/mt19937_64_points.json: operating points drawn by std::mt19937_64 and
// libstdc++'s std::uniform_real_distribution (Mach first, then the angle of attack in degrees).
// Build and run: g++ -O2 -o /tmp/pts mt19937_64_points.cpp && /tmp/pts > mt19937_64_points.json
#include <cstdio>
#include <random>

int main()
{
	const unsigned long long seeds[3] = { 4242ull, 2000ull, 200ull };
	const int counts[3] = { 64, 16, 16 };
	std::printf("[\n");
	for(int s = 0; s < 3; s++) {
		std::mt19937_64 g(seeds[s]);
		std::uniform_real_distribution<double> M(0.3, 0.75), A(-4.0, 6.0);
		double m[64], a[64];
		for(int i = 0; i < counts[s]; i++) { m[i] = M(g); a[i] = A(g); }
		std::printf("  {\"seed\": %llu, \"mach\": [", seeds[s]);
		for(int i = 0; i < counts[s]; i++) std::printf("%s%.17g", i ? ", " : "", m[i]);
		std::printf("], \"aoa_deg\": [");
		for(int i = 0; i < counts[s]; i++) std::printf("%s%.17g", i ? ", " : "", a[i]);
		std::printf("]}%s\n", s < 2 ? "," : "");
	}
	std::printf("]\n");
	return 0;
}
