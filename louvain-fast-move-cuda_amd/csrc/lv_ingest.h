// lv_ingest.h -- .edge text to a device-resident CSR on the GPU (lv_ingest.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <string>

#include "../../include/graph.h"
#include "lv_kernels.h"

namespace lv {

struct IngestTimes {
    double read_upload = 0;   // file read + host-to-device copy (overlapped)
    double tokenize = 0;      // reads of the text on the device
    double build = 0;         // normalise, sort, deduplicate, CSR
};

// The same graph GraphHost(path, weighted, skip_line) builds (graph.cu:17-87),
// built on the device; min_id receives the id shift.  Throws lv::Error.
GraphDevice ingest_edge_file(const std::string &path, bool weighted, int skip_line, hipStream_t s,
                             int64_t *min_id, IngestTimes *times);

}  // namespace lv
