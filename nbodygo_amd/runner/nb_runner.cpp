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
std::vector<Body> BodyCollection::TakeAdds() {
  std::lock_guard<std::mutex> g(lock_);
  std::vector<Body> a;
  a.swap(adds_);
  return a;
}
bool BodyCollection::Cycle(double R, const uint8_t *keep, const std::vector<Body> &adds,
                           const std::vector<Body> &frags) {
  std::lock_guard<std::mutex> g(lock_);
  bool changed = false;
  std::vector<Body> a;
  a.reserve(arr_.size() + frags.size() + adds.size());
  for (size_t i = 0; i < arr_.size(); i++) {
    if (keep ? keep[i] != 0 : arr_[i].exists)
      a.push_back(arr_[i]);
    else
      changed = true;
  }
  for (auto it = frags.rbegin(); it != frags.rend(); ++it) {  // newest first, as nb_cycle
    Body c = *it;
    c.r = R;
    a.push_back(c);
    changed = true;
  }
  for (auto it = adds.rbegin(); it != adds.rend(); ++it) {  // newest first, as nb_cycle
    Body c = *it;
    c.r = R;
    a.push_back(c);
    changed = true;
  }
  if (changed) arr_.swap(a);
  cycle_++;
  return changed;
}

// ---- runner -------------------------------------------------------------------------------------
ComputationRunner::ComputationRunner(int worker_cnt, double time_scaling, bool barnes_hut,
                                     ResultQueueHolder *rqh, BodyCollection *bc,
                                     const nb_params &params)
    : worker_cnt_(worker_cnt), time_scaling_(time_scaling), rqh_(rqh), bc_(bc) {
  (void)barnes_hut;
  events_ = params.events != 0;
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

nb_body ComputationRunner::ToDevice(const Body &b) {
  nb_body d{};
  d.x = b.x, d.y = b.y, d.z = b.z, d.vx = b.vx, d.vy = b.vy, d.vz = b.vz;
  d.m = b.mass, d.r = b.radius, d.rest = b.r, d.frag = b.frag_factor, d.frag_step = b.frag_step;
  d.beh = b.collision_behavior;
  d.flags = (uint8_t)((b.exists ? NB_F_EXISTS : 0) | (b.fragmenting ? NB_F_FRAGMENTING : 0) |
                      (b.is_sun ? NB_F_SUN : 0) | (b.pinned ? NB_F_PINNED : 0));
  return d;
}

void ComputationRunner::FromDevice(const nb_body &d, Body *b) {
  b->x = d.x, b->y = d.y, b->z = d.z, b->vx = d.vx, b->vy = d.vy, b->vz = d.vz;
  b->mass = d.m, b->radius = d.r, b->r = d.rest, b->frag_factor = d.frag, b->frag_step = d.frag_step;
  b->collision_behavior = d.beh;
  b->exists = (d.flags & NB_F_EXISTS) != 0;
  b->fragmenting = (d.flags & NB_F_FRAGMENTING) != 0;
  b->is_sun = (d.flags & NB_F_SUN) != 0;
  b->pinned = (d.flags & NB_F_PINNED) != 0;
}

int ComputationRunner::SyncToHost() {
  if (!host_stale_) return NB_OK;
  auto &arr = bc_->GetArray();
  const size_t n = arr.size();
  if (n == 0) {
    host_stale_ = false;
    return NB_OK;
  }
  std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n), m(n), r(n), rest(n), frag(n),
      fstep(n);
  std::vector<int32_t> beh(n);
  std::vector<uint8_t> fl(n);
  const int rc = nb_store_bodies(ctx_, x.data(), y.data(), z.data(), vx.data(), vy.data(),
                                 vz.data(), m.data(), r.data(), beh.data(), fl.data(),
                                 rest.data(), frag.data(), fstep.data());
  if (rc != NB_OK) {
    err_ = nb_last_error(ctx_);
    return rc;
  }
  for (size_t i = 0; i < n; i++) {
    nb_body d;
    d.x = x[i], d.y = y[i], d.z = z[i], d.vx = vx[i], d.vy = vy[i], d.vz = vz[i];
    d.m = m[i], d.r = r[i], d.rest = rest[i], d.frag = frag[i], d.beh = beh[i], d.flags = fl[i];
    d.frag_step = fstep[i];
    FromDevice(d, &arr[i]);
  }
  host_stale_ = false;
  return NB_OK;
}

// The one nb_load of a run (an empty collection loads too, so later adds have a context).
int ComputationRunner::pushIfDirty() {
  if (!dirty_) return NB_OK;
  auto &arr = bc_->GetArray();
  const size_t n = arr.size();
  std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n), m(n), r(n), rest(n), frag(n),
      fstep(n);
  std::vector<int32_t> beh(n);
  std::vector<uint8_t> fl(n);
  for (size_t i = 0; i < n; i++) {
    const nb_body d = ToDevice(arr[i]);
    x[i] = d.x, y[i] = d.y, z[i] = d.z, vx[i] = d.vx, vy[i] = d.vy, vz[i] = d.vz;
    m[i] = d.m, r[i] = d.r, rest[i] = d.rest, frag[i] = d.frag, fstep[i] = d.frag_step;
    beh[i] = d.beh;
    fl[i] = d.flags;
  }
  int rc = nb_load(ctx_, (int64_t)n, x.data(), y.data(), z.data(), vx.data(), vy.data(),
                   vz.data(), m.data(), r.data(), beh.data(), fl.data(), rest.data(), frag.data());
  if (rc == NB_OK && events_) rc = nb_set_frag_steps(ctx_, fstep.data());
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

// processDeletes (computation-runner.go:176-216) on the device: nb_remove_bodies walks the
// array as the reference does, closure-return quirk included (reaching the requested count
// returns from one callback only, so the walk goes on and the next stride point removes again).
void ComputationRunner::processDeletes() {
  const auto d = del_box_.Poll();
  if (!d || *d == 0) return;
  if (pushIfDirty() != NB_OK) return;
  int64_t removed = 0;
  if (!bc_->GetArray().empty() && nb_remove_bodies(ctx_, *d, &removed) != NB_OK) {
    err_ = nb_last_error(ctx_);
    std::fprintf(stderr, "[ERROR] computation runner: %s\n", err_.c_str());
    return;
  }
  if (removed) host_stale_ = true;
  std::fprintf(stderr, "[INFO] Computation runner set %lld bodies to not exist\n",
               (long long)removed);
}

void ComputationRunner::handleRequests() {
  std::deque<Request *> todo;
  {
    std::lock_guard<std::mutex> g(req_m_);
    todo.swap(requests_);
  }
  if (todo.empty()) return;
  const bool ready = pushIfDirty() == NB_OK;
  for (Request *r : todo) {
    if (ready) {
      auto &arr = bc_->GetArray();
      if (r->kind == 1) {  // HandleGetBody: one body read from the device
        for (size_t i = 0; i < arr.size(); i++) {
          const Body &b = arr[i];
          if ((!r->name.empty() && r->name == b.name) || r->id == b.id) {
            nb_body d;
            if (nb_get_body(ctx_, (int64_t)i, &d) != NB_OK) {
              err_ = nb_last_error(ctx_);
              break;
            }
            Body out = b;
            FromDevice(d, &out);
            r->body = out;
            break;
          }
        }
      } else {  // HandleModBody: each match read, modified and written back on the device
        int found = 0, modified = 0;
        for (size_t i = 0; i < arr.size(); i++) {
          Body &b = arr[i];
          if ((!r->cls.empty() && r->cls == b.cls) || (!r->name.empty() && r->name == b.name) ||
              r->id == b.id) {
            found++;
            nb_body d;
            if (nb_get_body(ctx_, (int64_t)i, &d) != NB_OK) {
              err_ = nb_last_error(ctx_);
              continue;
            }
            Body t = b;
            FromDevice(d, &t);
            if (!t.ApplyMods(r->mods)) continue;
            const nb_body nd = ToDevice(t);
            if (nb_set_body(ctx_, (int64_t)i, &nd) != NB_OK) {
              err_ = nb_last_error(ctx_);
              continue;
            }
            b = t;
            modified++;
          }
        }
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
  // Cycle: compaction and adds on the device, the host collection compacted alongside
  std::vector<Body> adds = bc_->TakeAdds();
  nb_event_stats evs{};
  if (events_) nb_get_event_stats(ctx_, &evs);
  if (gone || !adds.empty() || evs.fragments_pending > 0) {
    const size_t k = adds.size();
    std::vector<double> ax(k), ay(k), az(k), avx(k), avy(k), avz(k), am(k), ar(k), af(k), afs(k);
    std::vector<int32_t> abeh(k);
    std::vector<uint8_t> afl(k);
    for (size_t i = 0; i < k; i++) {
      const nb_body d = ToDevice(adds[i]);
      ax[i] = d.x, ay[i] = d.y, az[i] = d.z, avx[i] = d.vx, avy[i] = d.vy, avz[i] = d.vz;
      am[i] = d.m, ar[i] = d.r, af[i] = d.frag, afs[i] = d.frag_step, abeh[i] = d.beh;
      afl[i] = d.flags;
    }
    int64_t n_new = 0;
    rc = nb_add_bodies(ctx_, (int64_t)k, ax.data(), ay.data(), az.data(), avx.data(), avy.data(),
                       avz.data(), am.data(), ar.data(), abeh.data(), afl.data(), af.data(),
                       afs.data());
    if (rc == NB_OK) rc = nb_cycle(ctx_, R_, &n_new);
    if (rc != NB_OK) {
      err_ = nb_last_error(ctx_);
      return rc;
    }
    // the device's fragments (fragment(), fragcalc.go:97-117) come before the host's adds:
    // Elastic bodies of radius 1 with fresh ids, newest first
    std::vector<Body> frags((size_t)evs.fragments_pending);
    for (auto &f : frags) {
      f.id = NextId();
      f.radius = kFragmentRadius;
      f.collision_behavior = NB_CB_ELASTIC;
    }
    bc_->Cycle(R_, n ? ex_.data() : nullptr, adds, frags);
    if ((size_t)n_new != bc_->GetArray().size()) {
      err_ = "device and host collections disagree on the body count after Cycle";
      return NB_E_STATE;
    }
    host_stale_ = true;
  } else {
    bc_->Cycle(R_, nullptr, adds, {});
  }
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
