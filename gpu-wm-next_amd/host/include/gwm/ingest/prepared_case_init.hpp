// Reference include path include/gwm/ingest/prepared_case_init.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
