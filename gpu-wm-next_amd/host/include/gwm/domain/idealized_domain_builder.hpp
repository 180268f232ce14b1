// Reference include path include/gwm/domain/idealized_domain_builder.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
