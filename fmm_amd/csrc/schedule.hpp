// schedule.hpp -- host-side builders of the contraction schedules and leaf-kernel rows
// shared by the device operators.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "operator.hpp"

namespace fmmh {

// Builds one contraction schedule from target columns and their (group, source) lists.
// cols: output gid per column in task order, COLS per task (kNone = padding);
// groups_of(c): ascending (table, source gid) pairs of column c.
struct ScheduleBuilder {
  int width = dev::COLS;  // columns per task
  std::vector<uint32_t> task_group_off{0}, group_src, col_out;
  std::vector<uint16_t> group_table;
  uint64_t pairs = 0, slots = 0;

  explicit ScheduleBuilder(int w = dev::COLS) : width(w) {}

  // Groups are ordered by table, or by rank[table] when a rank is given.
  void add_task(const std::vector<uint32_t> &cols, const std::vector<std::vector<std::pair<uint16_t, uint32_t>>> &lists,
                const uint16_t *rank = nullptr) {
    const int C = width;
    std::map<uint16_t, std::pair<uint16_t, std::array<uint32_t, dev::COLS>>> by_key;
    for (int c = 0; c < C; ++c) {
      col_out.push_back(c < int(cols.size()) ? cols[c] : kNone);
      if (c >= int(cols.size())) continue;
      for (const auto &e : lists[c]) {
        const uint16_t key = rank ? rank[e.first] : e.first;
        auto it = by_key.find(key);
        if (it == by_key.end()) {
          std::array<uint32_t, dev::COLS> a;
          a.fill(kNone);
          it = by_key.emplace(key, std::make_pair(e.first, a)).first;
        }
        it->second.second[c] = e.second;
        ++pairs;
      }
    }
    for (const auto &kv : by_key) {
      group_table.push_back(kv.second.first);
      group_src.insert(group_src.end(), kv.second.second.begin(), kv.second.second.begin() + C);
      slots += C;
    }
    task_group_off.push_back(static_cast<uint32_t>(group_table.size()));
  }

  void append(const ScheduleBuilder &o) {
    const uint32_t base = static_cast<uint32_t>(group_table.size());
    for (size_t t = 1; t < o.task_group_off.size(); ++t) task_group_off.push_back(base + o.task_group_off[t]);
    group_src.insert(group_src.end(), o.group_src.begin(), o.group_src.end());
    col_out.insert(col_out.end(), o.col_out.begin(), o.col_out.end());
    group_table.insert(group_table.end(), o.group_table.begin(), o.group_table.end());
    pairs += o.pairs;
    slots += o.slots;
  }
  uint32_t ntasks() const { return static_cast<uint32_t>(task_group_off.size() - 1); }

  void finish(Schedule &s) {
    s.width = width;
    s.ntasks = static_cast<uint32_t>(task_group_off.size() - 1);
    s.pairs = pairs;
    s.slots = slots;
    s.task_group_off.upload(task_group_off);
    s.group_src.upload(group_src);
    s.col_out.upload(col_out);
    s.group_table.upload(group_table);
  }
};

inline void upload_rows(Rows &r, const std::vector<uint32_t> &id, const std::vector<uint32_t> &cell,
                 const std::vector<uint32_t> &off, const std::vector<uint32_t> &ent) {
  r.rows = static_cast<uint32_t>(id.size());
  r.row_id.upload(id);
  r.row_cell.upload(cell);
  r.row_off.upload(off);
  r.row_ent.upload(ent);
}


}  // namespace fmmh
