// Reference include path include/gwm/core/runtime_summary.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
