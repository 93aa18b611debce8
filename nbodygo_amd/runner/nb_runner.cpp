// nb_runner.cpp -- see nb_runner.hpp.
#include "nb_runner.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace nbr {

// ---- ids, mods --------------------------------------------------------------------------------
int64_t NextId() {
  static std::mutex m;
  static int64_t id = 0;
  std::lock_guard<std::mutex> g(m);
  return id++;
}

static std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}
static std::string upper(std::string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}
static std::string trim(const std::string &s) {  // strings.TrimSpace for ASCII white space
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

// strconv.ParseFloat(s, 64) for the decimal, hexadecimal (with a p exponent) and inf / nan forms;
// out-of-range values are errors, as ParseFloat's ErrRange is.
static bool go_parse_float(const std::string &s, double *out) {
  size_t i = 0, n = s.size();
  if (n == 0) return false;
  if (s[i] == '+' || s[i] == '-') i++;
  const std::string rest = lower(s.substr(i));
  if (rest == "inf" || rest == "infinity" || rest == "nan") {
    *out = std::strtod(s.c_str(), nullptr);
    return true;
  }
  bool hex = rest.size() > 2 && rest[0] == '0' && rest[1] == 'x';
  size_t j = i + (hex ? 2 : 0), digits = 0;
  auto isd = [&](char c) { return hex ? std::isxdigit((unsigned char)c) : std::isdigit((unsigned char)c); };
  while (j < n && isd(s[j])) j++, digits++;
  if (j < n && s[j] == '.') {
    j++;
    while (j < n && isd(s[j])) j++, digits++;
  }
  if (digits == 0) return false;
  const char e = hex ? 'p' : 'e';
  if (j < n && std::tolower((unsigned char)s[j]) == e) {
    j++;
    if (j < n && (s[j] == '+' || s[j] == '-')) j++;
    size_t ed = 0;
    while (j < n && std::isdigit((unsigned char)s[j])) j++, ed++;
    if (ed == 0) return false;
  } else if (hex) {
    return false;  // a hexadecimal mantissa needs its p exponent
  }
  if (j != n) return false;
  errno = 0;
  const double v = std::strtod(s.c_str(), nullptr);
  if (std::isinf(v)) return false;  // overflow
  *out = v;
  return true;
}

static bool go_parse_bool(const std::string &s, bool *out) {  // strconv.ParseBool
  static const char *t[] = {"1", "t", "T", "TRUE", "true", "True"};
  static const char *f[] = {"0", "f", "F", "FALSE", "false", "False"};
  for (auto *q : t)
    if (s == q) return *out = true, true;
  for (auto *q : f)
    if (s == q) return *out = false, true;
  return false;
}

static bool parse_boolean(const std::string &s) {  // globals.ParseBoolean
  const std::string l = lower(s);
  return l == "t" || l == "true" || l == "1" || l == "y" || l == "yes";
}

static const char *kCollision[] = {"none", "subsume", "elastic", "fragment"};
static const char *kColor[] = {"random", "black",   "white", "darkgray", "gray",
                               "lightgray", "red",  "green", "blue",     "yellow",
                               "magenta", "cyan",   "orange", "brown",   "pink"};

int32_t ParseCollisionBehavior(const std::string &s) {
  const std::string l = lower(s);
  for (int i = 0; i < 4; i++)
    if (l == kCollision[i]) return i;
  return NB_CB_ELASTIC;
}

int32_t ParseBodyColor(const std::string &s) {
  const std::string l = lower(s);
  for (int i = 0; i < 15; i++)
    if (l == kColor[i]) return i;
  return 0;
}

bool Body::ApplyMods(const std::vector<std::string> &mods) {
  for (const auto &mod : mods) {
    // strings.Split(mod, "="): exactly one '=' makes a name/value pair
    const size_t p = mod.find('=');
    if (p == std::string::npos || mod.find('=', p + 1) != std::string::npos) continue;
    const std::string k = upper(mod.substr(0, p)), v = mod.substr(p + 1);
    auto safe = [&](double cur) {
      double d;
      return go_parse_float(v, &d) ? d : cur;
    };
    if (k == "X") x = safe(x);
    else if (k == "Y") y = safe(y);
    else if (k == "Z") z = safe(z);
    else if (k == "VX") vx = safe(vx);
    else if (k == "VY") vy = safe(vy);
    else if (k == "VZ") vz = safe(vz);
    else if (k == "MASS") mass = safe(mass);
    else if (k == "RADIUS") radius = safe(radius);
    else if (k == "FRAG-FACTOR") frag_factor = safe(frag_factor);
    else if (k == "FRAG-STEP") frag_step = safe(frag_step);
    else if (k == "COLLISION") collision_behavior = ParseCollisionBehavior(v);
    else if (k == "COLOR") {
      if (!is_sun) color = ParseBodyColor(v);
    } else if (k == "EXISTS") exists = parse_boolean(v);
  }
  return true;
}

// ---- CSV ------------------------------------------------------------------------------------
// Records as encoding/csv reads them with FieldsPerRecord = -1 and Comment = '#': comma
// separated, optional double-quoted fields ("" inside is a quote, newlines allowed), empty lines
// and lines starting with '#' skipped, \r\n line ends.  A quote inside an unquoted field, or text
// after a closing quote, is a parse error.  Records go to `take` one at a time, as the reference
// reads them, until it returns false (so a malformed record after the last one wanted is never
// seen, as there).
static bool csv_records(const std::string &text,
                        const std::function<bool(const std::vector<std::string> &)> &take,
                        std::string *err) {
  size_t i = 0, n = text.size();
  int line = 1;
  while (i < n) {
    // start of a record
    if (text[i] == '\n') {
      i++, line++;
      continue;
    }
    if (text[i] == '\r' && i + 1 < n && text[i + 1] == '\n') {
      i += 2, line++;
      continue;
    }
    if (text[i] == '#') {
      while (i < n && text[i] != '\n') i++;
      continue;
    }
    std::vector<std::string> rec;
    for (;;) {
      std::string f;
      if (i < n && text[i] == '"') {
        i++;
        for (;;) {
          if (i >= n) {
            *err = "line " + std::to_string(line) + ": extraneous or missing \" in quoted-field";
            return false;
          }
          const char c = text[i];
          if (c == '"') {
            if (i + 1 < n && text[i + 1] == '"') {
              f += '"';
              i += 2;
              continue;
            }
            i++;
            break;
          }
          if (c == '\n') line++;
          if (c == '\r' && i + 1 < n && text[i + 1] == '\n') {
            i++;
            continue;
          }
          f += c;
          i++;
        }
        if (i < n && text[i] != ',' && text[i] != '\n' && text[i] != '\r') {
          *err = "line " + std::to_string(line) + ": extraneous or missing \" in quoted-field";
          return false;
        }
      } else {
        while (i < n && text[i] != ',' && text[i] != '\n') {
          if (text[i] == '"') {
            *err = "line " + std::to_string(line) + ": bare \" in non-quoted-field";
            return false;
          }
          if (text[i] == '\r' && (i + 1 == n || text[i + 1] == '\n')) {
            i++;
            continue;
          }
          f += text[i++];
        }
      }
      rec.push_back(f);
      if (i < n && text[i] == ',') {
        i++;
        continue;
      }
      if (i < n && text[i] == '\r') i++;
      if (i < n && text[i] == '\n') i++, line++;
      break;
    }
    if (!take(rec)) return true;
  }
  return true;
}

bool FromCsv(const std::string &path, int body_count, int32_t default_collision,
             int32_t default_color, std::vector<Body> *out, std::string *err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    *err = "Error opening csv: " + path;
    return false;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  int lines = 0;
  if (body_count <= 0) return true;
  return csv_records(ss.str(), [&](const std::vector<std::string> &f) -> bool {
    if (f.size() < 8) return true;  // index out of range: the row is skipped
    double v[8];
    bool ok = true;
    for (int k = 0; k < 8 && ok; k++) ok = go_parse_float(trim(f[k]), &v[k]);
    bool sun = false;
    if (ok && f.size() >= 9) ok = go_parse_bool(trim(f[8]), &sun);
    int32_t cb = default_collision, color = default_color;
    if (ok && f.size() >= 10) cb = ParseCollisionBehavior(trim(f[9]));
    if (ok && f.size() >= 11) color = ParseBodyColor(trim(f[10]));
    double ff = 0, fs = 0;
    if (ok && f.size() >= 12) ok = go_parse_float(trim(f[11]), &ff);
    if (ok && f.size() >= 13) ok = go_parse_float(trim(f[12]), &fs);
    if (!ok) return true;
    Body b;
    b.id = NextId();
    b.x = v[0], b.y = v[1], b.z = v[2], b.vx = v[3], b.vy = v[4], b.vz = v[5];
    b.mass = v[6], b.radius = v[7];
    b.collision_behavior = cb;
    b.color = color;
    b.frag_factor = ff;
    b.frag_step = fs;
    if (sun) b.SetSun(100);
    out->push_back(b);
    return ++lines < body_count;
  }, err);
}

bool WriteCsv(const std::string &path, const std::vector<Body> &bodies) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  for (const auto &b : bodies) {
    if (!b.exists) continue;
    std::fprintf(f, "%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%s,%s,%s,%.17g,%.17g\n", b.x,
                 b.y, b.z, b.vx, b.vy, b.vz, b.mass, b.radius, b.is_sun ? "true" : "false",
                 kCollision[std::min(std::max(b.collision_behavior, 0), 3)],
                 kColor[std::min(std::max(b.color, 0), 14)], b.frag_factor, b.frag_step);
  }
  return std::fclose(f) == 0;
}

// ---- result queues ----------------------------------------------------------------------------
std::optional<ResultQueue> ResultQueueHolder::NewResultQueue() {
  std::lock_guard<std::mutex> g(m_);
  if ((int)q_.size() >= max_) return std::nullopt;
  ResultQueue r;
  r.queue_num = queue_num_++;
  return r;
}
void ResultQueueHolder::Add(ResultQueue q) {
  std::lock_guard<std::mutex> g(m_);
  q_.push_back(std::move(q));
}
std::optional<ResultQueue> ResultQueueHolder::Next() {
  std::lock_guard<std::mutex> g(m_);
  if (q_.empty()) return std::nullopt;
  ResultQueue r = std::move(q_.front());
  q_.pop_front();
  return r;
}
int ResultQueueHolder::MaxQueues() {
  std::lock_guard<std::mutex> g(m_);
  return max_;
}
bool ResultQueueHolder::Resize(int max_queues) {
  std::lock_guard<std::mutex> g(m_);
  if (max_queues == max_) return false;
  max_ = max_queues;  // queued results are kept; new queues wait until the count falls below
  return true;
}
int ResultQueueHolder::Size() {
  std::lock_guard<std::mutex> g(m_);
  return (int)q_.size();
}

// ---- collection ---------------------------------------------------------------------------------
int BodyCollection::Count() {
  std::lock_guard<std::mutex> g(lock_);
  return (int)arr_.size();
}
void BodyCollection::EnqueueAdd(const Body &b) {
  std::lock_guard<std::mutex> g(lock_);
  adds_.push_back(b);
}
bool BodyCollection::Cycle(double R) {
  std::lock_guard<std::mutex> g(lock_);
  const size_t live = (size_t)std::count_if(arr_.begin(), arr_.end(),
                                            [](const Body &b) { return b.exists; });
  bool changed = false;
  if (live < arr_.size()) {
    std::vector<Body> a;
    a.reserve(live + adds_.size());
    for (auto &b : arr_)
      if (b.exists) a.push_back(b);
    arr_.swap(a);
    changed = true;
  }
  for (auto &b : adds_) {
    Body c = b;
    c.r = R;
    arr_.push_back(c);
    changed = true;
  }
  adds_.clear();
  cycle_++;
  return changed;
}

// ---- runner -------------------------------------------------------------------------------------
ComputationRunner::ComputationRunner(int worker_cnt, double time_scaling, bool barnes_hut,
                                     ResultQueueHolder *rqh, BodyCollection *bc,
                                     const nb_params &params)
    : worker_cnt_(worker_cnt), time_scaling_(time_scaling), rqh_(rqh), bc_(bc) {
  (void)barnes_hut;
  const int rc = nb_create(&params, &ctx_);
  if (rc != NB_OK) {
    err_ = std::string("nb_create failed: ") + nb_last_error(nullptr);
    ctx_ = nullptr;
  }
}

ComputationRunner::~ComputationRunner() {
  Stop();
  if (ctx_) nb_destroy(ctx_);
}

ComputationRunner &ComputationRunner::SetMaxIterations(int n) {
  max_iteration_ = n;
  return *this;
}

ComputationRunner &ComputationRunner::Start() {
  if (!running_ && ctx_) {
    stop_ = false;
    running_ = true;
    thread_ = std::thread([this] { run(); });
  }
  return *this;
}

void ComputationRunner::Stop() {
  stop_ = true;
  if (thread_.joinable()) thread_.join();
}

void ComputationRunner::Wait() {
  if (thread_.joinable()) thread_.join();
}

void ComputationRunner::run() {
  while (!stop_) {
    const int rc = RunOneComputation();
    if (rc != NB_OK && rc != NB_E_PAIR_OVERFLOW) {
      std::fprintf(stderr, "[ERROR] computation runner: %s\n", err_.c_str());
      break;
    }
    if (max_iteration_ > 0 && --max_iteration_ == 0) break;
  }
  // answer requests that arrived after the last cycle, then let later callers serve themselves
  handleRequests();
  {
    std::lock_guard<std::mutex> g(req_m_);
    running_ = false;
  }
  req_cv_.notify_all();
}

int ComputationRunner::SyncToHost() {
  if (!host_stale_) return NB_OK;
  auto &arr = bc_->GetArray();
  const size_t n = arr.size();
  std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n), rest(n);
  std::vector<uint8_t> fl(n);
  const int rc = nb_store(ctx_, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(),
                          fl.data(), rest.data());
  if (rc != NB_OK) {
    err_ = nb_last_error(ctx_);
    return rc;
  }
  for (size_t i = 0; i < n; i++) {
    Body &b = arr[i];
    b.x = x[i], b.y = y[i], b.z = z[i], b.vx = vx[i], b.vy = vy[i], b.vz = vz[i], b.r = rest[i];
    const bool ex = (fl[i] & NB_F_EXISTS) != 0;
    if (!ex && b.exists) b.SetNotExists();
    b.fragmenting = (fl[i] & NB_F_FRAGMENTING) != 0;
  }
  host_stale_ = false;
  return NB_OK;
}

int ComputationRunner::pushIfDirty() {
  if (!dirty_) return NB_OK;
  auto &arr = bc_->GetArray();
  const size_t n = arr.size();
  if (n == 0) {  // nothing to hold on the device; nb_step is not called for an empty collection
    dirty_ = false;
    host_stale_ = false;
    return NB_OK;
  }
  std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n), m(n), r(n), rest(n), frag(n);
  std::vector<int32_t> beh(n);
  std::vector<uint8_t> fl(n);
  for (size_t i = 0; i < n; i++) {
    const Body &b = arr[i];
    x[i] = b.x, y[i] = b.y, z[i] = b.z, vx[i] = b.vx, vy[i] = b.vy, vz[i] = b.vz;
    m[i] = b.mass, r[i] = b.radius, rest[i] = b.r, frag[i] = b.frag_factor;
    beh[i] = b.collision_behavior;
    fl[i] = (uint8_t)((b.exists ? NB_F_EXISTS : 0) | (b.fragmenting ? NB_F_FRAGMENTING : 0) |
                      (b.is_sun ? NB_F_SUN : 0) | (b.pinned ? NB_F_PINNED : 0));
  }
  const int rc = nb_load(ctx_, (int64_t)n, x.data(), y.data(), z.data(), vx.data(), vy.data(),
                         vz.data(), m.data(), r.data(), beh.data(), fl.data(), rest.data(),
                         frag.data());
  if (rc != NB_OK) {
    err_ = nb_last_error(ctx_);
    return rc;
  }
  dirty_ = false;
  host_stale_ = false;
  std::lock_guard<std::mutex> g(stats_m_);
  st_.loads++;
  return NB_OK;
}

// computation-runner.go:176-216, including its closure-return quirk: reaching the requested
// count returns from one callback only, so the walk goes on and the next stride point removes
// again.
void ComputationRunner::processDeletes() {
  const auto d = del_box_.Poll();
  if (!d || *d == 0) return;
  if (SyncToHost() != NB_OK) return;
  const int del_cnt = *d;
  int removed = 0;
  auto &arr = bc_->GetArray();
  if (del_cnt == -1) {
    for (auto &b : arr)
      if (b.exists) {
        b.SetNotExists();
        removed++;
      }
  } else if (del_cnt > 0) {
    const int count = (int)arr.size();
    const int step = del_cnt > count ? 1 : count / del_cnt;
    bool should_remove = false;
    int iter = 0;
    for (auto &b : arr) {
      if (iter % step == 0) should_remove = true;
      iter++;
      if (should_remove && !b.pinned && b.exists) {
        b.SetNotExists();
        should_remove = false;
        removed++;
      }
    }
  }
  if (removed) dirty_ = true;
  std::fprintf(stderr, "[INFO] Computation runner set %d bodies to not exist\n", removed);
}

void ComputationRunner::handleRequests() {
  std::deque<Request *> todo;
  {
    std::lock_guard<std::mutex> g(req_m_);
    todo.swap(requests_);
  }
  if (todo.empty()) return;
  const bool synced = SyncToHost() == NB_OK;
  for (Request *r : todo) {
    if (synced) {
      auto &arr = bc_->GetArray();
      if (r->kind == 1) {
        for (auto &b : arr)
          if ((!r->name.empty() && r->name == b.name) || r->id == b.id) {
            r->body = b;
            break;
          }
      } else {
        int found = 0, modified = 0;
        for (auto &b : arr)
          if ((!r->cls.empty() && r->cls == b.cls) || (!r->name.empty() && r->name == b.name) ||
              r->id == b.id) {
            found++;
            if (b.ApplyMods(r->mods)) modified++;
          }
        if (modified) dirty_ = true;
        r->result = found == 0 ? 0 : modified == 0 ? 1 : found == modified ? 3 : 2;
      }
    }
    std::lock_guard<std::mutex> g(req_m_);
    r->done = true;
  }
  req_cv_.notify_all();
}

std::optional<Body> ComputationRunner::GetBody(int64_t id, const std::string &name) {
  Request r;
  r.kind = 1;
  r.id = id;
  r.name = name;
  {
    std::lock_guard<std::mutex> g(req_m_);
    requests_.push_back(&r);
  }
  if (!running_) handleRequests();
  std::unique_lock<std::mutex> g(req_m_);
  req_cv_.wait(g, [&] { return r.done || !running_; });
  if (!r.done) {  // the runner stopped before answering
    g.unlock();
    handleRequests();
  }
  return r.body;
}

int ComputationRunner::ModBody(int64_t id, const std::string &name, const std::string &cls,
                               const std::vector<std::string> &mods) {
  Request r;
  r.kind = 2;
  r.id = id;
  r.name = name;
  r.cls = cls;
  r.mods = mods;
  {
    std::lock_guard<std::mutex> g(req_m_);
    requests_.push_back(&r);
  }
  if (!running_) handleRequests();
  std::unique_lock<std::mutex> g(req_m_);
  req_cv_.wait(g, [&] { return r.done || !running_; });
  if (!r.done) {
    g.unlock();
    handleRequests();
  }
  return r.result;
}

int ComputationRunner::RunOneComputation() {
  if (!ctx_) return NB_E_STATE;
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = cycle();
  std::lock_guard<std::mutex> g(stats_m_);
  st_.iterations++;
  st_.wall_ms +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int ComputationRunner::cycle() {
  if (auto t = time_scale_box_.Poll()) time_scaling_ = *t;  // updateTimeScaling
  if (auto r = R_box_.Poll()) R_ = *r;                      // updateCoefficientOfRestitution
  processDeletes();
  handleRequests();  // HandleGetBody, HandleModBody
  auto rq = rqh_->NewResultQueue();
  if (!rq) {  // the consumer is behind: skip this cycle (computation-runner.go:277-281)
    std::lock_guard<std::mutex> g(stats_m_);
    st_.skipped_no_queue++;
    return NB_OK;
  }
  int rc = pushIfDirty();
  if (rc != NB_OK) return rc;
  auto &arr = bc_->GetArray();
  const size_t n = arr.size();
  nb_stats s{};
  if (n > 0) {
    rc = nb_step(ctx_, time_scaling_, R_, 1, &s);
    if (rc != NB_OK && rc != NB_E_PAIR_OVERFLOW) {
      err_ = nb_last_error(ctx_);
      return rc;
    }
    host_stale_ = true;
    xyz_.resize(3 * n);
    ex_.resize(n);
    const int rc2 = nb_store_render(ctx_, xyz_.data(), ex_.data());
    if (rc2 != NB_OK) {
      err_ = nb_last_error(ctx_);
      return rc2;
    }
  }
  // Update's renderables (body.go:114-139 -> renderable.go:22-40)
  bool gone = false;
  for (size_t i = 0; i < n; i++) {
    const Body &b = arr[i];
    Renderable r;
    r.id = b.id;
    r.exists = ex_[i] != 0;
    if (r.exists) {
      r.x = xyz_[3 * i], r.y = xyz_[3 * i + 1], r.z = xyz_[3 * i + 2];
      r.radius = b.radius;
      r.is_sun = b.is_sun;
      r.intensity = (float)b.intensity;
      r.color = b.color;
    } else {
      gone = true;
    }
    rq->Add(r);
  }
  rqh_->Add(std::move(*rq));
  // Cycle: compaction and adds work on the host copy, so pull the advanced state first
  if (gone || bc_->HasAdds()) {
    rc = SyncToHost();
    if (rc != NB_OK) return rc;
  }
  if (bc_->Cycle(R_)) dirty_ = true;
  std::lock_guard<std::mutex> g(stats_m_);
  st_.computations++;
  st_.step_ms += s.wall_ms;
  st_.interactions += s.interactions;
  st_.pairs_total += s.pairs_last;
  st_.removed_total += s.removed_total;
  if (gone) st_.compactions++;
  return rc;
}

RunnerStats ComputationRunner::stats() const {
  std::lock_guard<std::mutex> g(stats_m_);
  return st_;
}

void ComputationRunner::PrintStats() const {
  const RunnerStats s = stats();
  const double fps = s.wall_ms > 0 ? s.computations / (s.wall_ms * 1e-3) : 0.0;
  std::printf(
      "Runner\n workerCnt: %d\n iterations: %llu\n computations: %llu\n skipped (no result "
      "queue): %llu\n device loads: %llu\n compactions: %llu\n step ms/computation: %.4f\n"
      " frames per second: %.3f\n interactions: %.6g\n interactions/s: %.6g\n elapsed ms: %.3f\n",
      worker_cnt_, (unsigned long long)s.iterations, (unsigned long long)s.computations,
      (unsigned long long)s.skipped_no_queue, (unsigned long long)s.loads,
      (unsigned long long)s.compactions, s.computations ? s.step_ms / s.computations : 0.0, fps,
      s.interactions, s.wall_ms > 0 ? s.interactions / (s.wall_ms * 1e-3) : 0.0, s.wall_ms);
}

}  // namespace nbr
