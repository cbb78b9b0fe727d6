// Drives libstdc++'s std::discrete_distribution with a scripted 32-bit
// engine so that each draw sees a chosen canonical uniform u, and prints the
// index it returns.  Input (stdin): n, n weights, m, then m pairs of 32-bit
// words (low, high) that generate_canonical<double, 53> combines into u.
// Output: for each pair, u (hex float) and the chosen index.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

struct Scripted {
    using result_type = uint32_t;
    std::vector<uint32_t> words;
    size_t pos = 0;
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return 0xFFFFFFFFu; }
    result_type operator()() { return words[pos++]; }
};

int main() {
    int n = 0;
    if (std::scanf("%d", &n) != 1) return 1;
    std::vector<double> w(n);
    for (auto& x : w) if (std::scanf("%lf", &x) != 1) return 1;
    int m = 0;
    if (std::scanf("%d", &m) != 1) return 1;
    Scripted eng;
    for (int i = 0; i < 2 * m; ++i) {
        unsigned long v = 0;
        if (std::scanf("%lu", &v) != 1) return 1;
        eng.words.push_back((uint32_t)v);
    }
    std::discrete_distribution<size_t> dist(w.begin(), w.end());
    for (int i = 0; i < m; ++i) {
        Scripted probe;
        probe.words = {eng.words[2 * i], eng.words[2 * i + 1]};
        double u = std::generate_canonical<double, 53>(probe);
        size_t k = dist(eng);
        std::printf("%a %zu\n", u, k);
    }
    return 0;
}
