int finess_run(int argc, char* argv[]);
int main(int argc, char* argv[]) { return finess_run(argc, argv); }
