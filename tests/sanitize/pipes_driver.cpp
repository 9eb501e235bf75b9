// Bounded-buffer pipes (include/tdvision/pipe.hpp) under ThreadSanitizer or
// AddressSanitizer: several producer/consumer chains with small buffers,
// WriteGuard finishing a pipe on unwind, readers seeing the end. Prints
// "pipes sanitize ok".
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "tdvision/pipe.hpp"

using namespace tdv;

int main()
{
    // a three-stage chain: producer -> doubler -> consumer, capacity 1 and 3
    for (int cap : {1, 3}) {
        ReadWritePipe<int> a(cap), b(cap);
        const int N = 20000;
        std::thread prod([&] {
            for (int i = 0; i < N; i++) a.write(i);
            a.finish();
        });
        std::thread mid([&] { // one element per update, as a work unit does
            for (;;) {
                WriteGuard<ReadWritePipe<int>> g(b);
                int v;
                if (a.read(&v)) g.write(2 * v);
                if (!g.wasWrite()) break; // the guard finished b
            }
        });
        long long sum = 0;
        int v, cnt = 0;
        while (b.read(&v)) {
            sum += v;
            cnt++;
        }
        prod.join();
        mid.join();
        if (cnt != N || sum != (long long)N * (N - 1)) return 1;
    }
    // a stage whose second update throws before writing: its guard
    // finishes the output on unwind and the reader ends after one element
    {
        ReadWritePipe<int> out(2);
        std::thread t([&] {
            try {
                for (int k = 0;; k++) {
                    WriteGuard<ReadWritePipe<int>> g(out);
                    if (k == 1) throw std::runtime_error("stage failed");
                    g.write(1);
                }
            } catch (const std::exception &) {
            }
        });
        int v, cnt = 0;
        while (out.read(&v)) cnt++;
        t.join();
        if (cnt != 1) return 2;
    }
    // many writers into one pipe
    {
        ReadWritePipe<int> p(4);
        std::vector<std::thread> ws;
        for (int w = 0; w < 4; w++)
            ws.emplace_back([&p, w] {
                for (int i = 0; i < 5000; i++) p.write(w);
            });
        std::thread closer([&] {
            for (auto &t : ws) t.join();
            p.finish();
        });
        int v, cnt = 0;
        while (p.read(&v)) cnt++;
        closer.join();
        if (cnt != 20000) return 3;
    }
    printf("pipes sanitize ok\n");
    return 0;
}
