// ORACLE -- test infrastructure only.  Rate-and-state friction restatement
// (filled in with the friction solver).
