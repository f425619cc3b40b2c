// schedule.hpp -- the schedule handle of include/gal_tree.h, shared by the
// host walk (trees.cpp) and the GPU executor (tree_exec.hip).  Internal: not
// part of the C ABI.
#pragma once
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/gal_tree.h"

struct gal_schedule {
  std::vector<gal_schedule_op> ops;
  int n_waves = 0;
  // Data that an executor derives from the operations once per device (the
  // GPU executor's operation graph), keyed by device and released with the
  // schedule through the deleter its creator supplies.
  using Attachment = std::unique_ptr<void, void (*)(void*)>;
  std::mutex attach_mutex;
  std::vector<std::pair<int, Attachment>> attachments;
};
