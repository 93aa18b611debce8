// nb_server.cpp -- command-line driver of the per-cycle path in C++, the counterpart of the
// reference server's --no-render mode (cmd/server/main.go): load bodies from a CSV, run the
// computation runner for a number of cycles or milliseconds, print the runner's statistics and
// optionally write the final state as CSV.
//
// Options (the reference's names where they exist, cmd/server/main.go:104-220):
//   -f, --csv PATH          body CSV (fromcsv.go schema)
//   -b, --bodies N          at most N bodies from the CSV (default 1000)
//   -c, --collision NAME    default collision behaviour for rows without one (default elastic)
//   -m, --scaling DT        time scaling per cycle (default 1e-9, the reference's default)
//   -t, --threads N         accepted for compatibility; the device replaces the worker pool
//   -u, --run-millis MS     stop after MS milliseconds
//   -r, --no-render         accepted (always headless)
//   --no-barnes-hut         accepted (the direct sum always runs)
//   --iterations N          stop after N cycles (the runner's SetMaxIterations)
//   --restitution R         coefficient of restitution (default 1)
//   --G G, --eps EPS        gravitational constant (default 6.673e-11) and Plummer softening
//   --no-cutoff             softened form without the d > Ri+Rj gate
//   --no-collisions         skip detection and resolution
//   --fp64                  FP64 force path (bit-identical to the reference's arithmetic)
//   --events                subsume and fragmentation events on the device (nb_params.events)
//   --remove N@K            before cycle K, ask the runner to remove N bodies (RemoveBodies)
//   --add PATH@K            before cycle K, enqueue every body of the CSV at PATH (AddBody);
//                           may repeat
//   --mod ID@K:M1,M2,...    before cycle K, ModBody on body id ID with "name=value" mods;
//                           may repeat
//   --get ID@K              before cycle K, print GetBody's answer for body id ID; may repeat
//   --out PATH              write the surviving bodies as CSV after the run
//   --parse-only            parse the CSV, print one line per body and exit (no device use)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "nb_runner.hpp"

static void usage() {
  std::fprintf(stderr,
               "usage: nb_server -f bodies.csv [-b N] [-c behaviour] [-m dt] [-u ms] "
               "[--iterations N]\n"
               "                 [--restitution R] [--G G] [--eps EPS] [--no-cutoff] "
               "[--no-collisions] [--fp64] [--events]\n"
               "                 [--remove N@K] [--add PATH@K] [--mod ID@K:M,...] [--get ID@K]\n"
               "                 [--out final.csv] [--parse-only]\n");
}

int main(int argc, char **argv) {
  std::string csv, out;
  int bodies = 1000, iterations = 0, run_millis = -1, remove_n = 0, remove_at = -1;
  double scaling = 1e-9, R = 1.0;
  int32_t collision = NB_CB_ELASTIC;
  bool parse_only = false;
  struct Action {
    int at;
    int kind;  // 0 add, 1 mod, 2 get
    std::string path;
    int64_t id = -1;
    std::vector<std::string> mods;
  };
  std::vector<Action> actions;
  auto split_at = [](const std::string &s, std::string *head, int *k) {
    const size_t at = s.find('@');
    *head = s.substr(0, at);
    *k = at == std::string::npos ? 0 : std::atoi(s.substr(at + 1).c_str());
  };
  nb_params p;
  nb_params_init(&p);
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i], v;
    const size_t eq = a.find('=');
    if (a.rfind("--", 0) == 0 && eq != std::string::npos) {  // --opt=value
      v = a.substr(eq + 1);
      a = a.substr(0, eq);
    }
    auto val = [&]() -> std::string {
      if (!v.empty()) return v;
      if (i + 1 >= argc) {
        usage();
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "-f" || a == "--csv") csv = val();
    else if (a == "-b" || a == "--bodies") bodies = std::atoi(val().c_str());
    else if (a == "-c" || a == "--collision") collision = nbr::ParseCollisionBehavior(val());
    else if (a == "-m" || a == "--scaling") scaling = std::atof(val().c_str());
    else if (a == "-t" || a == "--threads") (void)val();
    else if (a == "-u" || a == "--run-millis") run_millis = std::atoi(val().c_str());
    else if (a == "-r" || a == "--no-render" || a == "--no-barnes-hut") {
    } else if (a == "--iterations") iterations = std::atoi(val().c_str());
    else if (a == "--restitution") R = std::atof(val().c_str());
    else if (a == "--G") p.G = std::atof(val().c_str());
    else if (a == "--eps") p.eps = std::atof(val().c_str());
    else if (a == "--no-cutoff") p.overlap_cutoff = 0;
    else if (a == "--no-collisions") p.collisions = 0;
    else if (a == "--fp64") p.precision = NB_PREC_FP64;
    else if (a == "--events") p.events = 1;
    else if (a == "--remove") {
      const std::string s = val();
      const size_t at = s.find('@');
      remove_n = std::atoi(s.substr(0, at).c_str());
      remove_at = at == std::string::npos ? 0 : std::atoi(s.substr(at + 1).c_str());
    } else if (a == "--add") {
      Action ac;
      ac.kind = 0;
      split_at(val(), &ac.path, &ac.at);
      actions.push_back(ac);
    } else if (a == "--mod") {
      const std::string s = val();
      const size_t colon = s.find(':');
      Action ac;
      ac.kind = 1;
      std::string head;
      split_at(s.substr(0, colon), &head, &ac.at);
      ac.id = std::atoll(head.c_str());
      std::string rest = colon == std::string::npos ? "" : s.substr(colon + 1);
      size_t q = 0;
      while (q <= rest.size() && !rest.empty()) {
        const size_t c = rest.find(',', q);
        ac.mods.push_back(rest.substr(q, c == std::string::npos ? std::string::npos : c - q));
        if (c == std::string::npos) break;
        q = c + 1;
      }
      actions.push_back(ac);
    } else if (a == "--get") {
      Action ac;
      ac.kind = 2;
      std::string head;
      split_at(val(), &head, &ac.at);
      ac.id = std::atoll(head.c_str());
      actions.push_back(ac);
    } else if (a == "--out") out = val();
    else if (a == "--parse-only") parse_only = true;
    else if (a == "-h" || a == "--help") {
      usage();
      return 0;
    } else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      usage();
      return 2;
    }
  }
  if (csv.empty()) {
    usage();
    return 2;
  }
  std::vector<nbr::Body> bs;
  std::string err;
  if (!nbr::FromCsv(csv, bodies, collision, 0, &bs, &err)) {
    std::fprintf(stderr, "Sim CSV Reader failed reading CSV. The error is: %s\n", err.c_str());
    return 1;
  }
  if (parse_only) {
    for (const auto &b : bs)
      std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %d %d %.17g %.17g\n", b.x,
                  b.y, b.z, b.vx, b.vy, b.vz, b.mass, b.radius, b.is_sun ? 1 : 0,
                  b.collision_behavior, b.color, b.frag_factor, b.frag_step);
    return 0;
  }
  nbr::BodyCollection bc(std::move(bs));
  nbr::ResultQueueHolder rqh(1 << 30);  // headless: the queues are drained below
  nbr::ComputationRunner runner(1, scaling, false, &rqh, &bc, p);
  if (!runner.ok()) {
    std::fprintf(stderr, "%s\n", runner.error().c_str());
    return 1;
  }
  runner.SetCoefficientOfRestitution(R);
  const auto t0 = std::chrono::steady_clock::now();
  int rc = NB_OK;
  for (int it = 0; iterations <= 0 || it < iterations; it++) {
    if (run_millis >= 0 &&
        std::chrono::steady_clock::now() - t0 >= std::chrono::milliseconds(run_millis))
      break;
    if (it == remove_at && remove_n != 0) runner.RemoveBodies(remove_n);
    for (const auto &ac : actions) {
      if (ac.at != it) continue;
      if (ac.kind == 0) {
        std::vector<nbr::Body> add;
        if (!nbr::FromCsv(ac.path, 1 << 30, collision, 0, &add, &err)) {
          std::fprintf(stderr, "add: %s\n", err.c_str());
          return 1;
        }
        for (const auto &b : add) runner.AddBody(b);
      } else if (ac.kind == 1) {
        std::printf("mod %lld: %d\n", (long long)ac.id, runner.ModBody(ac.id, "", "", ac.mods));
      } else {
        const auto b = runner.GetBody(ac.id, "");
        if (b)
          std::printf("get %lld: %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d\n",
                      (long long)ac.id, b->x, b->y, b->z, b->vx, b->vy, b->vz, b->mass,
                      b->radius, b->exists ? 1 : 0);
        else
          std::printf("get %lld: none\n", (long long)ac.id);
      }
    }
    rc = runner.RunOneComputation();
    while (rqh.Next()) {
    }  // the renderer's side: consume the results
    if (rc != NB_OK && rc != NB_E_PAIR_OVERFLOW) {
      std::fprintf(stderr, "cycle %d failed: %s\n", it, runner.error().c_str());
      return 1;
    }
    if (iterations <= 0 && run_millis < 0) break;  // neither limit: one cycle
  }
  runner.PrintStats();
  if (!out.empty()) {
    if (runner.SyncToHost() != NB_OK || !nbr::WriteCsv(out, bc.GetArray())) {
      std::fprintf(stderr, "could not write %s\n", out.c_str());
      return 1;
    }
  }
  return 0;
}
