// main.cpp -- the ltpmg command-line driver (the reference's src/main.cpp).
#include "common.h"
#include "database.h"
#include "scheduler.h"

int main(int argc, char **argv) {
    initialize_dependency(argc, argv);
    auto scheduler = std::make_shared<Scheduler>();
    scheduler->start();
    scheduler->end();
    free_dependency();
    return 0;
}
