// Checks the engine's host thread pool (sapporo2_amd/csrc/hostpool.h) on the
// CPU: every element of every job is visited exactly once, jobs of all sizes
// (inline and pooled), two caller threads at once (one of them runs inline
// while the other holds the pool), workers that sleep between jobs, and a
// forked child (which has none of the workers and runs its jobs itself).
#include <sys/wait.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "hostpool.h"

static int check(unsigned seed, int reps) {
  unsigned s = seed;
  for (int r = 0; r < reps; ++r) {
    s = s * 1103515245u + 12345u;
    const size_t n = (s >> 8) % 70000;
    std::vector<int> hits(n, 0);
    sap::par_for(n, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) hits[i] += 1;
    });
    for (size_t i = 0; i < n; ++i)
      if (hits[i] != 1) {
        std::printf("rep %d n %zu: element %zu visited %d times\n", r, n, i, hits[i]);
        return 1;
      }
  }
  return 0;
}

int main(int argc, char** argv) {
  const int reps = argc > 1 ? std::atoi(argv[1]) : 300;
  int bad = check(1, reps);
  int bad2 = 0, bad3 = 0;
  std::thread a([&] { bad2 = check(7, reps); });
  std::thread b([&] { bad3 = check(11, reps); });
  a.join();
  b.join();
  int bad4 = 0;
  if (argc > 2 && std::strcmp(argv[2], "fork") == 0) {
    const pid_t p = fork();
    if (p == 0) _exit(check(13, 20));
    int st = 0;
    bad4 = (p < 0 || waitpid(p, &st, 0) != p || !WIFEXITED(st) || WEXITSTATUS(st) != 0) ? 1 : 0;
    bad4 |= check(17, 20);  // the parent's pool still works
  }
  const int bad_all = bad | bad2 | bad3 | bad4;
  std::printf("threads %u: %s\n", sap::HostPool::get().threads(), bad_all ? "FAIL" : "ok");
  return bad_all;
}
