// Definitions behind the opaque handles of include/gfaz_amd.h.
#pragma once

#include "gfaz_amd.h"

#include "codec.hpp"
#include "ctx.hpp"
#include "resident.hpp"

#include <memory>

struct gfz_ctx {
  std::unique_ptr<gfz::Ctx> ctx;
};

struct gfz_rulebook {
  gfz_ctx *owner = nullptr;
  gfz::Rulebook rb;
};

struct gfz_resident {
  gfz_ctx *owner = nullptr;
  const gfz_rulebook *rulebook = nullptr;
  gfz::Resident res;
};
