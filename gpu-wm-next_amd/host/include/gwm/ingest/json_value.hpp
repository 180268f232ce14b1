// Reference include path include/gwm/ingest/json_value.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
