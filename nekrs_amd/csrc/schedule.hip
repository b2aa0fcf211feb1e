// schedule.hip -- MULTIGRID SCHEDULE (the .par key pMGSchedule): the rules of
// parseMultigridSchedule (src/solvers/elliptic/MG/parseMultigridSchedule.cpp:7-131),
// the level orders determineMGLevels takes from it (MG/determineMGLevels.cpp:22-40)
// and the per-level Chebyshev degrees the level setup reads from it
// (MG/ellipticMultiGridLevelSetup.cpp:160-167).  Host code only.
#include <algorithm>
#include <cctype>
#include <climits>
#include <cstdlib>
#include <map>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "schedule.hpp"

namespace nrs {

namespace {

// the pieces of `in` between `dlim`s after all white space is removed, empty
// pieces dropped (serializeString, src/core/nrssys.hpp:148-160)
std::vector<std::string> pieces(const std::string &in, char dlim)
{
  std::string s;
  for (char c : in)
    if (!std::isspace((unsigned char)c)) s += c;
  std::vector<std::string> out;
  size_t a = 0;
  while (a <= s.size()) {
    size_t b = s.find(dlim, a);
    if (b == std::string::npos) b = s.size();
    if (b > a) out.push_back(s.substr(a, b - a));
    a = b + 1;
  }
  return out;
}

// std::stoi of the second '='-piece of a token: an optional sign and at least
// one digit at its start, anything after them ignored
bool value_of(const std::string &token, int &v)
{
  const auto p = pieces(token, '=');
  if (p.size() < 2) return false;
  const char *c = p[1].c_str();
  char *end = nullptr;
  const long x = std::strtol(c, &end, 10);
  if (end == c || x < INT_MIN || x > INT_MAX) return false;
  v = (int)x;
  return true;
}

}  // namespace

void parse_mg_schedule(const std::string &schedule, int N, int defaultDegree, nrs_elliptic_opts *o)
{
  const int INVALID = INT_MIN;
  std::string err;
  const auto entries = pieces(schedule, ',');

  // the coarsest order first: entries of that order need no degree
  int minOrder = INT_MAX;
  for (const auto &entry : entries)
    for (const auto &token : pieces(entry, '+'))
      if (token.find('p') != std::string::npos) {
        int v;
        if (!value_of(token, v)) throw Error(NRS_ERR_INVALID, "schedule: no order in token '" + token + "'");
        minOrder = std::min(minOrder, v);
      }

  // (order, down leg) -> degree; the down leg ends at the first increase of the order
  std::map<std::pair<int, bool>, int> sched;
  std::vector<int> downOrders;
  bool downLeg = true;
  int prevOrder = INT_MAX;
  for (const auto &entry : entries) {
    int order = INVALID, degree = INVALID;
    for (const auto &token : pieces(entry, '+')) {
      int v = 0;
      if (token.find('p') != std::string::npos) {
        value_of(token, v);
        order = v;
        if (order > prevOrder) downLeg = false;
        if (downLeg) downOrders.push_back(order);
        prevOrder = order;
      } else if (token.find("degree") != std::string::npos) {
        if (!value_of(token, v)) throw Error(NRS_ERR_INVALID, "schedule: no value in token '" + token + "'");
        degree = v;
      } else {
        err += "unknown token '" + token + "' in schedule '" + schedule + "'!\n";
      }
    }
    if (order == INVALID) {
      err += "order not specified in " + entry + "\n";
      continue;
    }
    if (degree == INVALID && order != minOrder) degree = defaultDegree;
    sched[{order, downLeg}] = degree;
  }

  std::set<int> down, up;
  for (const auto &kv : sched) {
    if (kv.first.first <= minOrder) continue;
    (kv.first.second ? down : up).insert(kv.first.first);
  }
  if (down != up) err += "down leg and up leg orders must be identical!\n";
  for (const auto &kv : sched)
    if (kv.first.first != minOrder && kv.second == INVALID)
      err += "degree not specified for order " + std::to_string(kv.first.first) + "!\n";
  for (int l : downOrders)
    if (l < 1) err += "encountered level = " + std::to_string(l) + " < 1!\n";
  if (downOrders.empty())
    err += "no levels in schedule '" + schedule + "'!\n";
  else if (downOrders.front() != N)
    err += "finest degree " + std::to_string(downOrders.front()) + " does not match polynomial degree " +
           std::to_string(N) + "!\n";
  // A degree on the coarsest order is refused: the coarsest level is solved or
  // smoothed by its own settings.  (The reference's level setup would take it as
  // the coarsest level's degree after prolongation and, through std::map's
  // operator[], set the one before restriction to 0.)
  for (const auto &kv : sched)
    if (kv.first.first == minOrder && kv.second != INVALID)
      err += "degree " + std::to_string(kv.second) + " given for the coarsest order " +
             std::to_string(minOrder) + ", which keeps its own smoothing!\n";
  for (const auto &kv : sched)
    if (kv.first.first != minOrder && kv.second != INVALID && (kv.second < 1 || kv.second > 16))
      err += "degree " + std::to_string(kv.second) + " of order " + std::to_string(kv.first.first) +
             " outside 1..16!\n";

  // levels: the distinct down-leg orders, finest first (determineMGLevels.cpp:31-38)
  std::vector<int> levels;
  for (const auto &kv : sched)
    if (kv.first.second) levels.push_back(kv.first.first);
  std::sort(levels.rbegin(), levels.rend());
  if (levels.size() > 8) err += "more than 8 levels\n";
  if (!err.empty()) throw Error(NRS_ERR_INVALID, "schedule: " + err);

  o->nLevels = (int)levels.size();
  for (int i = 0; i < 8; ++i) {
    o->levels[i] = i < (int)levels.size() ? levels[i] : 0;
    o->preDegree[i] = 0;
    o->postDegree[i] = 0;
  }
  // ellipticMultiGridLevelSetup.cpp:163-166: UpLegChebyshevDegree (after
  // prolongation) <- the down-leg entry, DownLegChebyshevDegree (before
  // restriction) <- the up-leg entry.  The coarsest level keeps its own.
  for (int i = 0; i + 1 < (int)levels.size(); ++i) {
    o->postDegree[i] = sched.at({levels[i], true});
    o->preDegree[i] = sched.at({levels[i], false});
  }
}

}  // namespace nrs
